"""Readers for MGmol's text inputs and a box configuration built from them.

* ``read_cfg``: the Boost.program_options configuration file (``mgmol.cfg``:
  ``key=value`` lines, ``[Section]`` headers, ``#`` comments), as
  Control / read_config parse it (src/read_config.cc).
* ``read_lrs``: localization centres, one ``x y z`` line per orbital
  (the ``lrs.in`` of tests/EnergyAndForces, read by LocalizationRegions).
* ``read_coords``: atoms, ``name species x y z`` per line (``coords.in``).
* ``box_setup``: per-orbital boxes around the centres on the configuration's
  mesh, with MGmol's Gaussian initial orbitals (LocGridOrbitals::initGauss,
  src/LocGridOrbitals.cc:553-630: exp(-r^2 / rc^2), minimum image, cut at
  6 rc, rc = Orbitals.initial_width), normalised.  The local potential is a
  stand-in: Gaussian wells at the atoms (no pseudopotentials here).
"""
import numpy as np

from .synth import BoxConfig


def _value(s):
    for conv in (int, float):
        try:
            return conv(s)
        except ValueError:
            pass
    return s


def read_cfg(path):
    """{section: {key: value}}; keys before the first header go to section ''."""
    out = {"": {}}
    sec = ""
    with open(path) as f:
        for raw in f:
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            if line.startswith("[") and line.endswith("]"):
                sec = line[1:-1].strip()
                out.setdefault(sec, {})
                continue
            if "=" not in line:
                raise ValueError("%s: cannot parse %r" % (path, raw))
            k, v = line.split("=", 1)
            out[sec][k.strip()] = _value(v.strip())
    return out


# Control::setOptions, src/Control.cc:1353-1359: FDtype name -> lap_type
FD_TYPES = {"Mehrstellen": 0, "2nd": 1, "4th": 2, "6th": 3, "8th": 4, "Mehrstellen2": 10}


def fd_type(cfg_path):
    """The orbital discretisation of an MGmol config (default Mehrstellen,
    src/read_config.cc:52) as a lap_type code."""
    name = str(read_cfg(cfg_path)[""].get("FDtype", "Mehrstellen"))
    if name not in FD_TYPES:
        raise ValueError("%s: unknown FDtype %r" % (cfg_path, name))
    return FD_TYPES[name]


def read_lrs(path):
    """(n, 3) centres in bohr."""
    a = np.loadtxt(path, ndmin=2)
    if a.shape[1] < 3:
        raise ValueError("%s: fewer than 3 columns" % path)
    return a[:, :3].astype(np.float64)


def read_coords(path):
    """[(name, species, (x, y, z))] in bohr."""
    out = []
    with open(path) as f:
        for raw in f:
            t = raw.split("#", 1)[0].split()
            if not t:
                continue
            out.append((t[0], int(t[1]), tuple(float(x) for x in t[2:5])))
    return out


def _mesh(cfg):
    n = np.array([cfg["Mesh"][k] for k in ("nx", "ny", "nz")], dtype=np.int64)
    o = np.array([float(cfg["Domain"][k]) for k in ("ox", "oy", "oz")])
    ll = np.array([float(cfg["Domain"][k]) for k in ("lx", "ly", "lz")])
    h = ll / n
    if not np.allclose(h, h[0], rtol=1e-12):
        raise ValueError("the box solver takes one spacing for all three axes")
    return n, o, ll, float(h[0])


def box_setup(cfg_path, lrs_path, coords_path=None, edge=28, levels=3, well_depth=-1.0, well_width=1.0):
    """BoxConfig for MGmol inputs: one box of ``edge`` points per
    localization centre, its origin the centre's grid point minus edge/2,
    rounded down to a multiple of 2^(levels-1) and wrapped into the grid."""
    cfg = read_cfg(cfg_path)
    n, o, ll, h = _mesh(cfg)
    cen = read_lrs(lrs_path)
    rc = float(cfg.get("Orbitals", {}).get("initial_width", 1.0))
    m = 1 << (levels - 1)
    if edge % m or 2 * edge + 4 > min(n):
        raise ValueError("box edge %d does not fit the %s mesh with %d levels" % (edge, tuple(n), levels))
    nb = len(cen)
    origin = np.zeros((nb, 3), dtype=np.int32)
    for d in range(3):
        start = np.floor((cen[:, d] - o[d]) / h).astype(np.int64) - edge // 2
        origin[:, d] = np.mod((start // m) * m, n[d])
    edges = np.full(nb, edge, dtype=np.int32)
    off = np.zeros(nb + 1, dtype=np.int64)
    off[1:] = np.cumsum(edges.astype(np.int64) ** 3)
    psi = np.zeros(int(off[-1]))
    vel = h ** 3
    for j in range(nb):
        r2 = 0.0
        for d in range(3):
            x = o[d] + (origin[j, d] + np.arange(edge)) * h
            dd = x - cen[j, d]
            dd -= ll[d] * np.round(dd / ll[d])
            sh = [1, 1, 1]
            sh[d] = edge
            r2 = r2 + (dd * dd).reshape(sh)
        val = np.where(r2 < (6.0 * rc) ** 2, np.exp(-r2 / (rc * rc)), 0.0)
        val /= np.sqrt(vel * np.sum(val * val))
        psi[off[j]:off[j + 1]] = val.reshape(-1)
    v = np.zeros(tuple(int(x) for x in n))
    atoms = read_coords(coords_path) if coords_path else []
    for _, _, pos in atoms:
        r2 = 0.0
        for d in range(3):
            x = o[d] + np.arange(n[d]) * h
            dd = x - pos[d]
            dd -= ll[d] * np.round(dd / ll[d])
            sh = [1, 1, 1]
            sh[d] = int(n[d])
            r2 = r2 + (dd * dd).reshape(sh)
        v += well_depth * np.exp(-r2 / (2.0 * well_width ** 2))
    return BoxConfig(tuple(int(x) for x in n), h, origin, edges, off, psi, 2.0 * v, cen, levels,
                     {"cfg": cfg, "atoms": atoms, "initial_width": rc, "domain_origin": o})
