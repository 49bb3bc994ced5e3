"""Property tests of the host-side parts of the library (no GPU): the celldump
encoder/decoder (write_files.f90 / read_hydro.f90 layout) on random grids, and
the EOS table builder (eos.f90:127-520) over random mixtures."""
import dataclasses

import numpy as np
from hypothesis import given, settings
from hypothesis import strategies as st


@settings(max_examples=25, deadline=None)
@given(nx=st.integers(5, 9), ny=st.integers(5, 9), nz=st.integers(1, 8), seed=st.integers(0, 2**31 - 1),
       step=st.integers(0, 10**7), time=st.floats(0.0, 1e6, allow_nan=False))
def test_celldump_round_trip(tmp_path_factory, nx, ny, nz, seed, step, time):
    """decode(encode(state)) gives back rho, E, u, p exactly, the momenta as
    u*rho, adindx = gammafix and the frame cells reset (read_hydro.f90:150-178),
    and the file has the record structure 1234 / time,step / units / cells."""
    from boxzyhydro_amd.hydro import celldump_decode, celldump_encode
    from boxzyhydro_amd.ics import boundary_flags
    from boxzyhydro_amd.params import HydroParams

    prm = HydroParams(nx=nx, ny=ny, nz=nz, dx=0.3, dy=0.2, dz=0.1, gammafix=1.4)
    n = prm.ngrid
    rng = np.random.Generator(np.random.PCG64(seed))
    cons = rng.uniform(0.1, 2.0, (n, 5))
    u = rng.normal(size=(n, 3))
    p = rng.uniform(0.1, 2.0, n)
    path = str(tmp_path_factory.mktemp("cd") / "celldump")
    celldump_encode(prm, path, time, step, cons, u, p)
    raw = open(path, "rb").read()
    assert len(raw) == (4 + 4 + 4) + (4 + 12 + 4) + (4 + 64 + 4) + n * (4 + 92 + 4)
    assert np.frombuffer(raw[:12], "<i4").tolist() == [4, 1234, 4]
    t2, s2, out = celldump_decode(prm, path)
    assert t2 == time and s2 == step
    frame = boundary_flags(prm).reshape(-1) == 1
    inner = ~frame
    assert np.array_equal(out["cons"][inner, 0], cons[inner, 0])
    assert np.array_equal(out["cons"][inner, 4], cons[inner, 4])
    assert np.array_equal(out["u"], u) and np.array_equal(out["p"], p)
    assert np.array_equal(out["cons"][inner, 1:4], u[inner] * cons[inner, 0:1])
    assert np.all(out["cons"][frame] == [prm.small_rho, 0.0, 0.0, 0.0, prm.small_eps])
    assert np.all(out["adindx"] == prm.gammafix)


@settings(max_examples=12, deadline=None)
@given(h2=st.sampled_from([0, 1]), xabun=st.floats(0.5, 0.9), zabun=st.floats(0.0, 0.05),
       ac=st.floats(0.5, 3.0), bc=st.floats(0.5, 3.0), nrho=st.integers(2, 9),
       dtk=st.sampled_from([50.0, 125.0, 400.0]), tbg=st.floats(3.0, 40.0))
def test_eos_tables_finite_and_monotone(h2, xabun, zabun, ac, bc, nrho, dtk, tbg):
    """Over random mixtures and table sizes: every table finite, specific
    energy increasing with T in every density row, the log-energy grid of table
    2 non-decreasing, and the interpolated temperatures of table 2 inside the
    temperature range of table 1."""
    from boxzyhydro_amd.hydro import eos_tables
    from boxzyhydro_amd.params import box_params

    yabun = 1.0 - xabun - zabun
    prm = dataclasses.replace(box_params(8), H2STAT=h2, xabun=xabun, yabun=yabun, zabun=zabun, ac=ac, bc=bc,
                              NEOS_RHO=nrho, dTk_eos=dtk, tk_bgrnd=tbg)
    t = eos_tables(prm)
    for k in ("eng_table", "gamma_table", "muc_table", "eng_table2", "tk_table2", "muc_table2", "gamma_table2"):
        assert np.all(np.isfinite(t[k])), k
    assert np.all(np.diff(t["eng_table"], axis=1) > 0)
    assert np.all(np.diff(t["eng_table2"], axis=1) >= 0)
    tk = t["tk_table"]
    assert np.all(t["tk_table2"] >= tk[0] * (1 - 1e-12)) and np.all(t["tk_table2"] <= tk[-1] * (1 + 1e-12))
