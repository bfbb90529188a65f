"""Restart primitives and raw Tecplot files of device fields: files the
reference's routines wrote (tests/golden/io) import into device fields bit
for bit; device exports equal the host writer's files byte for byte and the
reference's value for value; a time step restarted from exported state
continues bit-identically."""
import os

import numpy as np
import pytest

import moons_amd as M

from test_io import GOLD, RAW_CASES, bits, raw_case, same_tokens

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,pad", RAW_CASES)
def test_raw_files_through_device_fields(gpu, tmp_path, name, pad):
    print(f"raw {name} pad={pad}", flush=True)
    hn, f, d, fname = raw_case(name, pad)
    m = gpu.mesh(*hn)
    u = gpu.SF(m, name).upload(f)
    gpu.export_raw(m, u, str(tmp_path) + "/dev_", "U", pad)
    gpu.raw_write(hn, name, f, str(tmp_path) + "/host_", "U", pad)
    dev, host = open(tmp_path / ("dev_" + fname)).read(), open(tmp_path / ("host_" + fname)).read()
    assert dev == host
    assert same_tokens(dev, open(d + fname).read())
    v = gpu.SF(m, name).fill(-7.0)
    gpu.import_raw(m, v, d, "U", pad)
    g = v.download()
    inner = tuple(slice(pad, s - pad) for s in f.shape)
    np.testing.assert_array_equal(bits(g[inner]), bits(f[inner]))
    mask = np.ones(f.shape, bool)
    mask[inner] = False
    assert np.all(g[mask] == -7.0)


def test_restart_primitives_of_device_fields(gpu, tmp_path):
    print("restart primitives 32^3", flush=True)
    m = gpu.mesh.cube_uniform(32)
    rng = np.random.default_rng(3)
    for dl in ("CC", "Fy", "Ez", "N"):
        u = gpu.SF(m, dl)
        a = rng.standard_normal(u.shape) * 10.0 ** rng.integers(-30, 30, u.shape)
        u.upload(a)
        path = str(tmp_path / f"u_{dl}.txt")
        u.export(path)
        np.testing.assert_array_equal(bits(gpu.gf_read(path)), bits(a))
        v = gpu.SF(m, dl).fill(0.0).import_(path)
        np.testing.assert_array_equal(bits(v.download()), bits(a))
    # the reference's file of another shape is refused, not read partially
    with pytest.raises(gpu.MoonsError, match="shape"):
        gpu.SF(m, "CC").import_(os.path.join(GOLD, "gf_reference.txt"))
    # a staggered VF exports one file per component
    U = [gpu.SF(m, 1 << d).fill(float(d)) for d in range(3)]
    gpu.export_raw(m, U, str(tmp_path) + "/", "U", 1, t=0.25)
    names = sorted(x for x in os.listdir(tmp_path) if x.endswith(".dat"))
    assert names == [f"Uf_{c}_t=0.25000000000000000000.dat" for c in "xyz"]


def test_step_restarts_bit_identically(gpu, tmp_path):
    """Low-Rem MHD cavity, method 6: three uninterrupted steps equal two steps,
    an export of the state (U, p, F, B, phi, FB), an import into fresh fields
    and a fresh step object, and one more step -- bit for bit."""
    from test_gpu_step import build

    print("restart mhd_m6", flush=True)
    isp = (5, 1e-6, 1e-13, 1)
    N = (10, 9, 8)
    g, m, st, ost, Ug, pg, kw = build(gpu, N, True, 1.0, isp)
    g2, m2, st2, ost2, Ug2, pg2, kw2 = build(gpu, N, True, 1.0, isp)
    st.step(3)
    st2.step(2)
    state = {"U": Ug2, "p": [pg2], "F": st2.field("F"), "B": kw2["B"], "phi": [kw2["phi"]], "FB": st2.field("FB")}
    for key, fields in state.items():
        for q, f in enumerate(fields):
            f.export(str(tmp_path / f"{key}{q}.txt"))
    del st2
    # fresh fields with the same BCs (the step object built with them is
    # dropped), filled from the files before a new step object is created
    g3, m3, st3, ost3, Ug3, pg3, kw3 = build(gpu, N, True, 1.0, isp)
    del st3
    for q in range(3):
        Ug3[q].import_(str(tmp_path / f"U{q}.txt"))
        kw3["B"][q].import_(str(tmp_path / f"B{q}.txt"))
    pg3.import_(str(tmp_path / "p0.txt"))
    kw3["phi"].import_(str(tmp_path / "phi0.txt"))
    st3 = gpu.MHD_step(m3, Ug3, pg3, **kw3)
    for key in ("F", "FB"):
        for q, f in enumerate(st3.field(key)):
            f.import_(str(tmp_path / f"{key}{q}.txt"))
    st3.step(1)
    for a, b in [(Ug, Ug3), ([pg], [pg3]), (kw["B"], kw3["B"]), ([kw["phi"]], [kw3["phi"]]),
                 (st.field("F"), st3.field("F")), (st.field("J"), st3.field("J"))]:
        for x, y in zip(a, b):
            np.testing.assert_array_equal(bits(x.download()), bits(y.download()))
