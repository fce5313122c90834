"""Writes colormaps.npz: the 256 listed RGB colours of the four colormaps the
benchmark scenes draw from (viridis, plasma, inferno, magma), taken from the
installed matplotlib.  They pin ColorTransferFunction.from_colormap, which
must reproduce a 256-colour listed map exactly with lut_size=256."""

import os

import matplotlib
import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "colormaps.npz")

if __name__ == "__main__":
    maps = {}
    for name in ("viridis", "plasma", "inferno", "magma"):
        cmap = matplotlib.colormaps.get_cmap(name)
        maps[name] = np.asarray(cmap.colors, dtype=np.float64)[:, :3]
    maps["matplotlib_version"] = np.array(matplotlib.__version__)
    np.savez(OUT, **maps)
    print(OUT, {k: v.shape for k, v in maps.items()})
