"""MGmol input readers on the reference's own test inputs
(tests/EnergyAndForces: mgmol.cfg, lrs.in, coords.in, kept as fixtures under
tests/golden/energy_and_forces) and the box configuration built from them."""
import os

import numpy as np
import pytest

from mgmol_amd import inputs

D = os.path.join(os.path.dirname(__file__), "golden", "energy_and_forces")
CFG, LRS, COORDS = (os.path.join(D, f) for f in ("mgmol.cfg", "lrs.in", "coords.in"))


def test_read_cfg():
    c = inputs.read_cfg(CFG)
    assert c[""]["FDtype"] == "Mehrstellen" and c[""]["verbosity"] == 1
    assert (c["Mesh"]["nx"], c["Mesh"]["ny"], c["Mesh"]["nz"]) == (64, 64, 64)
    assert c["Domain"]["ox"] == -6.0 and c["Domain"]["lz"] == 12.0
    assert c["Orbitals"]["initial_width"] == 1.5 and c["Quench"]["atol"] == 1e-9


def test_read_lrs_and_coords():
    cen = inputs.read_lrs(LRS)
    assert cen.shape == (6, 3)
    assert np.allclose(cen[0], [0.0, 0.7, 1.4]) and np.allclose(cen[5], [-0.46, 0.0, 0.0])
    atoms = inputs.read_coords(COORDS)
    assert [a[0] for a in atoms] == ["N1", "N2"] and atoms[1][2] == (0.0, 0.0, 1.0345)


def test_box_setup():
    cfg = inputs.box_setup(CFG, LRS, COORDS, edge=28, levels=3)
    assert cfg.gdim == (64, 64, 64) and cfg.h == pytest.approx(12.0 / 64)
    assert cfg.nbox == 6 and np.all(cfg.origin % 4 == 0)
    vel = cfg.h ** 3
    o = np.array([-6.0, -6.0, -6.0])
    for j in range(cfg.nbox):
        blk = cfg.psi[cfg.offset[j]:cfg.offset[j + 1]].reshape(28, 28, 28)
        assert vel * np.sum(blk * blk) == pytest.approx(1.0, rel=1e-12)
        # the maximum sits at the grid point nearest the centre (the centre is inside the box)
        imax = np.array(np.unravel_index(np.argmax(blk), blk.shape))
        g = (cfg.origin[j] + imax) % 64
        x = o + g * cfg.h
        assert np.all(np.abs(x - cfg.centers[j]) <= cfg.h)
    # wells of -1 Ha (width 1 bohr) at the two atoms 2.069 bohr apart, in Rydberg:
    # at an atom 2 (-1 - exp(-2.069^2 / 2)) = -2.47 Ry; the grid misses the atoms
    assert -2.47 < cfg.vtot.min() < -2.2 and cfg.vtot.max() <= 0.0


def test_box_setup_rejects_oversized_boxes():
    with pytest.raises(ValueError):
        inputs.box_setup(CFG, LRS, edge=32)


def test_fd_type_of_the_reference_inputs():
    from mgmol_amd import inputs
    import os
    d = os.path.join(os.path.dirname(__file__), "golden", "energy_and_forces")
    assert inputs.fd_type(os.path.join(d, "mgmol.cfg")) == 0  # FDtype=Mehrstellen
