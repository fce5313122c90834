"""Camera placement from spherical parameters, with a self-contained FP64
quaternion implementation.

``get_camera_pos`` reproduces the reference's placement rule
(``pyvr/camera/control.py:238-351``): the initial offset ``init_pos - target``
(rescaled to ``distance``) and ``init_up`` are rotated by
``R = R_azimuth * R_elevation * R_roll``, where

* ``R_azimuth`` turns by ``azimuth`` about ``init_up``,
* ``R_elevation`` turns by ``elevation`` about ``normalize(init_up x view_dir)``
  (``view_dir = -normalize(offset)``; a fallback axis is used when that cross
  product vanishes),
* ``R_roll`` turns by ``roll`` about ``view_dir``,

and ``A * B`` means "apply B first".  The reference builds these with
``scipy.spatial.transform.Rotation``; here they are unit quaternions in
float64.  The rotation vectors are formed with the same numpy expressions as
the reference (e.g. ``azimuth * init_up`` with a float32 ``init_up``), so the
float32 rounding of the rotation angles matches.

``CameraPath`` (keyframe interpolation) and ``CameraController`` (orbit,
zoom, arcball trackball, pan, roll, presets, transitions) follow the
reference's classes (``pyvr/camera/control.py:377-740``); they generate the
camera sequences of request streams (``pyvr_amd.hip_renderer.serve``).
"""

from typing import List, Optional, Tuple

import numpy as np

from .camera import Camera, _cross3, _norm, validate_camera_angles


def _quat_from_rotvec(v) -> np.ndarray:
    """Unit quaternion (x, y, z, w) for a rotation of |v| radians about v/|v|."""
    v = np.asarray(v, dtype=np.float64)
    theta = float(np.sqrt(v @ v))
    if theta < 1e-3:
        # sin(t/2)/t by its Taylor series near 0 (avoids 0/0 at the identity)
        t2 = theta * theta
        s = 0.5 - t2 / 48.0 + t2 * t2 / 3840.0
    else:
        s = np.sin(0.5 * theta) / theta
    q = np.array([v[0] * s, v[1] * s, v[2] * s, np.cos(0.5 * theta)])
    return q / np.sqrt(q @ q)


def _quat_mul(p: np.ndarray, q: np.ndarray) -> np.ndarray:
    """Hamilton product p*q of (x, y, z, w) quaternions: rotate by q, then by p."""
    pv, pw = p[:3], p[3]
    qv, qw = q[:3], q[3]
    v = pw * qv + qw * pv + _cross3(pv, qv)
    return np.array([v[0], v[1], v[2], pw * qw - pv @ qv])


def _quat_rotate(q: np.ndarray, x) -> np.ndarray:
    """Rotate vector x by unit quaternion q (via its 3x3 matrix, in float64)."""
    qx, qy, qz, qw = q
    m = np.array(
        [
            [1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw)],
            [2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw)],
            [2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)],
        ]
    )
    return m @ np.asarray(x, dtype=np.float64)


def get_camera_pos(
    target: np.ndarray,
    azimuth: float,
    elevation: float,
    roll: float,
    distance: float,
    init_pos: Optional[np.ndarray] = None,
    init_up: Optional[np.ndarray] = None,
) -> Tuple[np.ndarray, np.ndarray]:
    """Camera position and up vector (both float64) for spherical parameters.

    Defaults and errors follow the reference (control.py:289-320):
    ``init_pos`` defaults to ``[0, 0, distance]``, ``init_up`` to ``[0, 1, 0]``;
    ``CameraError`` for non-numeric/non-finite angles, ``ValueError`` for a bad
    target, distance or initial vectors.
    """
    validate_camera_angles(azimuth, elevation, roll)
    if not isinstance(target, np.ndarray) or target.shape != (3,):
        raise ValueError("target must be a 3D numpy array")
    if not isinstance(distance, (int, float)) or distance <= 0:
        raise ValueError("distance must be positive")
    if init_pos is None:
        init_pos = np.array([0, 0, distance], dtype=np.float32)
    if init_up is None:
        init_up = np.array([0, 1, 0], dtype=np.float32)
    if not isinstance(init_pos, np.ndarray) or init_pos.shape != (3,):
        raise ValueError("init_pos must be a 3D numpy array")
    if not isinstance(init_up, np.ndarray) or init_up.shape != (3,):
        raise ValueError("init_up must be a 3D numpy array")

    target = target.astype(np.float32)
    init_pos = init_pos.astype(np.float32)
    init_up = init_up.astype(np.float32)

    offset = init_pos - target
    norm = _norm(offset)
    if norm == 0:
        raise ValueError("init_pos must not be the zero vector (relative to target)")
    offset = offset / norm * distance

    q_az = _quat_from_rotvec(azimuth * init_up)

    view_dir = -offset / _norm(offset)
    axis = _cross3(init_up, view_dir)
    axis_norm = _norm(axis)
    if axis_norm < 1e-6:
        # init_up parallel to the view direction: pick a fixed helper axis
        helper = np.array([1, 0, 0]) if abs(init_up[0]) < 0.9 else np.array([0, 0, 1])
        axis = _cross3(init_up, helper)
        axis_norm = _norm(axis)
    axis = axis / axis_norm
    q_el = _quat_from_rotvec(elevation * axis)
    q_roll = _quat_from_rotvec(roll * view_dir)

    q = _quat_mul(_quat_mul(q_az, q_el), q_roll)
    position = _quat_rotate(q, offset) + target
    up = _quat_rotate(q, init_up)
    return position, up


def get_camera_pos_from_params(params: Camera) -> Tuple[np.ndarray, np.ndarray]:
    """``get_camera_pos`` on the fields of a ``Camera``."""
    return get_camera_pos(
        target=params.target,
        azimuth=params.azimuth,
        elevation=params.elevation,
        roll=params.roll,
        distance=params.distance,
        init_pos=params.init_pos,
        init_up=params.init_up,
    )


# ---------------------------------------------------------------------------
# Trackball helpers and camera animation (reference control.py:17-236, 377-740)

def _map_to_sphere(x: float, y: float, radius: float = 1.0) -> np.ndarray:
    """Arcball mapping of a normalised 2-D point onto a sphere of ``radius``
    (Shoemake's arcball, with Bell's hyperbolic sheet z = (r^2/2)/|p| beyond
    |p| = r/sqrt(2) so that the surface stays smooth), returned as a unit
    float32 vector; (0, 0) maps to (0, 0, 1)."""
    d2 = x * x + y * y
    r2 = radius * radius
    z = np.sqrt(r2 - d2) if d2 <= 0.5 * r2 else 0.5 * r2 / np.sqrt(d2)
    p = np.array([x, y, z], dtype=np.float32)
    return p / np.linalg.norm(p)


def _camera_basis(camera: Camera) -> np.ndarray:
    """3x3 float64 matrix whose columns are the camera's right, up and back
    (from target towards the camera) unit vectors, built from its placement:
    the up vector is re-orthogonalised against the view direction."""
    position, up = get_camera_pos_from_params(camera)
    forward = np.asarray(camera.target, dtype=np.float64) - position
    forward /= np.linalg.norm(forward)
    right = np.cross(forward, up)
    right /= np.linalg.norm(right)
    return np.column_stack([right, np.cross(right, forward), -forward])


def _camera_to_quaternion(camera: Camera):
    """The camera's orientation (columns right, up, back) as a
    ``scipy.spatial.transform.Rotation``, as the reference returns it."""
    from scipy.spatial.transform import Rotation

    return Rotation.from_matrix(_camera_basis(camera))


def _signed_angle(a: np.ndarray, b: np.ndarray, axis: np.ndarray) -> float:
    """Angle from a to b in (-pi, pi], positive counter-clockwise about axis."""
    return float(np.arctan2(np.dot(np.cross(a, b), axis), np.dot(a, b)))


def _quaternion_to_camera_angles(rotation, target: np.ndarray, distance: float, init_pos: np.ndarray,
                                 init_up: np.ndarray) -> Tuple[float, float, float]:
    """(azimuth, elevation, roll) that ``get_camera_pos`` turns into the
    orientation ``rotation`` (a scipy ``Rotation`` or a 3x3 matrix with
    columns right, up, back) about ``target`` at ``distance``.

    * azimuth: signed angle about ``init_up`` from the initial offset to the
      camera's offset, both projected onto the plane normal to ``init_up``;
      0 when either projection vanishes (camera at a pole);
    * elevation: asin of the back vector's component along ``init_up``;
    * roll: signed angle about the view direction from the up vector of
      (azimuth, elevation, roll = 0) to the orientation's up vector, both
      projected onto the plane normal to the view direction.
    """
    m = rotation.as_matrix() if hasattr(rotation, "as_matrix") else np.asarray(rotation, dtype=np.float64)
    up_vec, back = m[:, 1], m[:, 2]
    forward = -back / np.linalg.norm(back)
    target = np.asarray(target, dtype=np.float64)
    init_up = np.asarray(init_up, dtype=np.float64)
    offset = back * distance                        # camera position - target
    start = np.asarray(init_pos, dtype=np.float64) - target
    start = start / np.linalg.norm(start) * distance
    a = start - np.dot(start, init_up) * init_up
    b = offset - np.dot(offset, init_up) * init_up
    if np.linalg.norm(a) < 1e-9 or np.linalg.norm(b) < 1e-9:
        azimuth = 0.0
    else:
        azimuth = _signed_angle(a, b, init_up)
    elevation = float(np.arcsin(np.clip(-np.dot(forward, init_up), -1.0, 1.0)))
    _, plain_up = get_camera_pos(target=target.astype(np.float32), azimuth=azimuth, elevation=elevation,
                                 roll=0.0, distance=distance, init_pos=np.asarray(init_pos),
                                 init_up=np.asarray(init_up))
    e = plain_up - np.dot(plain_up, forward) * forward
    u = up_vec - np.dot(up_vec, forward) * forward
    if np.linalg.norm(e) < 1e-9 or np.linalg.norm(u) < 1e-9:
        roll = 0.0
    else:
        roll = _signed_angle(e, u, forward)
    return azimuth, elevation, roll


def _axis_angle_matrix(axis: np.ndarray, angle: float) -> np.ndarray:
    """Rotation matrix of ``angle`` about the unit ``axis`` (Rodrigues)."""
    k = np.array([[0.0, -axis[2], axis[1]], [axis[2], 0.0, -axis[0]], [-axis[1], axis[0], 0.0]])
    return np.eye(3) + np.sin(angle) * k + (1.0 - np.cos(angle)) * (k @ k)


class CameraPath:
    """Keyframed camera animation (reference control.py:377-489): piecewise
    linear in every parameter over equal-length segments, angles along the
    shorter arc."""

    def __init__(self, keyframes: List[Camera]):
        if len(keyframes) < 2:
            raise ValueError("At least 2 keyframes are required for a camera path")
        self.keyframes = keyframes
        self.validate_keyframes()

    def validate_keyframes(self) -> None:
        for i, k in enumerate(self.keyframes):
            if not isinstance(k, Camera):
                raise ValueError(f"Keyframe {i} must be a Camera instance")
            k.validate()

    @staticmethod
    def _lerp_vector(v1: np.ndarray, v2: np.ndarray, t: float) -> np.ndarray:
        return v1 + (v2 - v1) * t

    @staticmethod
    def _lerp_angle(a1: float, a2: float, t: float) -> float:
        """From a1 towards a2 along the shorter arc, both first wrapped to
        [-pi, pi)."""
        a1 = ((a1 + np.pi) % (2 * np.pi)) - np.pi
        a2 = ((a2 + np.pi) % (2 * np.pi)) - np.pi
        d = a2 - a1
        if d > np.pi:
            d -= 2 * np.pi
        elif d < -np.pi:
            d += 2 * np.pi
        return a1 + d * t

    def interpolate(self, t: float) -> Camera:
        """The camera at t in [0, 1] (0: first keyframe, 1: last)."""
        if not 0.0 <= t <= 1.0:
            raise ValueError("t must be between 0.0 and 1.0")
        if t == 0.0:
            return self.keyframes[0].copy()
        if t == 1.0:
            return self.keyframes[-1].copy()
        n = len(self.keyframes) - 1
        seg = min(int(t * n), n - 1)
        u = t * n - seg
        a, b = self.keyframes[seg], self.keyframes[seg + 1]
        lin = lambda x, y: x + (y - x) * u
        return Camera(
            target=self._lerp_vector(a.target, b.target, u),
            azimuth=self._lerp_angle(a.azimuth, b.azimuth, u),
            elevation=self._lerp_angle(a.elevation, b.elevation, u),
            roll=self._lerp_angle(a.roll, b.roll, u),
            distance=lin(a.distance, b.distance),
            init_pos=self._lerp_vector(a.init_pos, b.init_pos, u),
            init_up=self._lerp_vector(a.init_up, b.init_up, u),
            fov=lin(a.fov, b.fov),
            near_plane=lin(a.near_plane, b.near_plane),
            far_plane=lin(a.far_plane, b.far_plane),
        )

    def generate_frames(self, n_frames: int) -> List[Camera]:
        """``n_frames`` cameras at t = i / (n_frames - 1)."""
        if n_frames < 1:
            raise ValueError("n_frames must be at least 1")
        if n_frames == 1:
            return [self.keyframes[0].copy()]
        return [self.interpolate(i / (n_frames - 1)) for i in range(n_frames)]


class CameraController:
    """Interactive camera manipulation (reference control.py:492-740):
    orbit, zoom, arcball trackball, pan, roll, presets and transitions."""

    def __init__(self, initial_params: Optional[Camera] = None):
        self.params = (initial_params if initial_params is not None else Camera.front_view()).copy()

    def orbit(self, delta_azimuth: float, delta_elevation: float) -> None:
        """Turn about the target; elevation stays in [-pi/2, pi/2]."""
        self.params.azimuth += delta_azimuth
        self.params.elevation = float(np.clip(self.params.elevation + delta_elevation, -np.pi / 2, np.pi / 2))

    def zoom(self, factor: float) -> None:
        """Scale the distance (< 1 closer), kept in [0.01, 1000]."""
        if factor <= 0:
            raise ValueError("Zoom factor must be positive")
        self.params.distance = float(np.clip(self.params.distance * factor, 0.01, 1000.0))

    def trackball(self, dx: float, dy: float, viewport_width: int, viewport_height: int,
                  sensitivity: float = 1.0) -> None:
        """Arcball rotation for a mouse drag of (dx, dy) pixels (y down).

        The drag, scaled by the smaller viewport side and ``sensitivity`` and
        negated (dragging right turns the camera left), moves a point from the
        arcball's centre; the rotation taking (0, 0, 1) to its sphere point is
        applied in the camera's own frame, and the new orientation is
        decomposed back into azimuth, elevation and roll (distance and target
        unchanged).  Drags below 0.001 of the viewport are ignored."""
        if viewport_width <= 0 or viewport_height <= 0:
            raise ValueError(f"Viewport dimensions must be positive, got width={viewport_width}, "
                             f"height={viewport_height}")
        if dx == 0 and dy == 0:
            return
        scale = min(viewport_width, viewport_height)
        x = -(dx / scale) * sensitivity
        y = -(dy / scale) * sensitivity
        if abs(x) < 0.001 and abs(y) < 0.001:
            return
        p0 = _map_to_sphere(0.0, 0.0).astype(np.float64)
        p1 = _map_to_sphere(x, y).astype(np.float64)
        axis = np.cross(p0, p1)
        n = np.linalg.norm(axis)
        if n < 1e-8:
            return
        angle = float(np.arccos(np.clip(np.dot(p0, p1), -1.0, 1.0)))
        turned = _camera_basis(self.params) @ _axis_angle_matrix(axis / n, angle)
        az, el, roll = _quaternion_to_camera_angles(turned, self.params.target, self.params.distance,
                                                    self.params.init_pos, self.params.init_up)
        self.params.azimuth, self.params.elevation, self.params.roll = az, el, roll

    def pan(self, delta_target: np.ndarray) -> None:
        if not isinstance(delta_target, np.ndarray) or delta_target.shape != (3,):
            raise ValueError("delta_target must be a 3D numpy array")
        self.params.target += delta_target           # in place: the target keeps its array and dtype

    def roll_camera(self, delta_roll: float) -> None:
        self.params.roll += delta_roll

    def set_distance(self, distance: float) -> None:
        if distance <= 0:
            raise ValueError("Distance must be positive")
        self.params.distance = distance

    def look_at(self, target: np.ndarray) -> None:
        if not isinstance(target, np.ndarray) or target.shape != (3,):
            raise ValueError("target must be a 3D numpy array")
        self.params.target = target.astype(np.float32)

    _PRESETS = {"front": "front_view", "side": "side_view", "top": "top_view", "isometric": "isometric_view"}

    def reset_to_preset(self, preset: str, distance: Optional[float] = None) -> None:
        """One of 'front', 'side', 'top', 'isometric' about the current
        target, at ``distance`` or the current distance."""
        if preset not in self._PRESETS:
            raise ValueError(f"Unknown preset: {preset}. Available: {list(self._PRESETS)}")
        make = getattr(Camera, self._PRESETS[preset])
        self.params = make(target=self.params.target.copy(),
                           distance=self.params.distance if distance is None else distance)

    def get_position_and_up(self) -> Tuple[np.ndarray, np.ndarray]:
        return get_camera_pos_from_params(self.params)

    def animate_to(self, target_params: Camera, n_frames: int = 30) -> List[Camera]:
        """``n_frames`` cameras from the current one to ``target_params``."""
        return CameraPath([self.params.copy(), target_params.copy()]).generate_frames(n_frames)
