"""Writes colormaps.npz: 256 RGB colours of each of the 12 colormaps the
reference's colour selector offers (pyvr/interface/widgets.py:312-325), taken
from the installed matplotlib -- the listed colours of the listed maps
(viridis, plasma, inferno, magma, cividis, turbo) and the map sampled at
linspace(0, 1, 256) for the segmented ones (gray, bone, copper, hot, cool,
jet), which for a 256-entry map are its table entries.  They pin
ColorTransferFunction.from_colormap, which must reproduce them exactly with
lut_size=256; the c3 requests draw from the first four."""

import os

import matplotlib
import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "colormaps.npz")
UI_COLORMAPS = ["viridis", "plasma", "inferno", "magma", "cividis", "gray", "bone", "copper", "hot", "cool",
                "turbo", "jet"]

if __name__ == "__main__":
    maps = {}
    for name in UI_COLORMAPS:
        cmap = matplotlib.colormaps.get_cmap(name)
        if hasattr(cmap, "colors"):
            maps[name] = np.asarray(cmap.colors, dtype=np.float64)[:, :3]
        else:
            maps[name] = np.asarray(cmap(np.linspace(0.0, 1.0, 256)), dtype=np.float64)[:, :3]
        assert maps[name].shape == (256, 3), (name, maps[name].shape)
    maps["matplotlib_version"] = np.array(matplotlib.__version__)
    np.savez(OUT, **maps)
    print(OUT, {k: v.shape for k, v in maps.items()})
