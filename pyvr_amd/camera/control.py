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
"""

from typing import Optional, Tuple

import numpy as np

from .camera import Camera, validate_camera_angles


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
    v = pw * qv + qw * pv + np.cross(pv, qv)
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
    norm = np.linalg.norm(offset)
    if norm == 0:
        raise ValueError("init_pos must not be the zero vector (relative to target)")
    offset = offset / norm * distance

    q_az = _quat_from_rotvec(azimuth * init_up)

    view_dir = -offset / np.linalg.norm(offset)
    axis = np.cross(init_up, view_dir)
    axis_norm = np.linalg.norm(axis)
    if axis_norm < 1e-6:
        # init_up parallel to the view direction: pick a fixed helper axis
        helper = np.array([1, 0, 0]) if abs(init_up[0]) < 0.9 else np.array([0, 0, 1])
        axis = np.cross(init_up, helper)
        axis_norm = np.linalg.norm(axis)
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
