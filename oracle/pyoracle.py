"""ctypes wrapper of the C oracle (build/libboxzy_oracle.so) and runner for the
reference driver (_ref/ref_harness).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py.  The product path never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libboxzy_oracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")


class _Params(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
        ("gammafix", ctypes.c_double),
        ("small_rho", ctypes.c_double), ("small_eps", ctypes.c_double),
        ("vlimit", ctypes.c_double), ("cfl", ctypes.c_double),
        ("xabun", ctypes.c_double), ("yabun", ctypes.c_double),
        ("zabun", ctypes.c_double), ("mu_z", ctypes.c_double),
        ("avmag", ctypes.c_double * 3),
        ("no_outflow", ctypes.c_int * 6),
    ]


_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.bxo_create.restype = vp
        L.bxo_create.argtypes = [ctypes.POINTER(_Params)]
        L.bxo_destroy.argtypes = [vp]
        L.bxo_bind.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.bxo_ngrid.argtypes = [vp]
        L.bxo_nghost.argtypes = [vp]
        L.bxo_geometry.argtypes = [vp, ip, ip]
        L.bxo_muc.argtypes = [vp]
        L.bxo_muc.restype = ctypes.c_double
        L.bxo_rgas.argtypes = [vp]
        L.bxo_rgas.restype = ctypes.c_double
        for f in ("bxo_state", "bxo_velocity", "bxo_set_ghost_cells", "bxo_copy_old", "bxo_avisc"):
            getattr(L, f).argtypes = [vp]
        L.bxo_courant.argtypes = [vp]
        L.bxo_courant.restype = ctypes.c_double
        L.bxo_source.argtypes = [vp, ctypes.c_double]
        L.bxo_flux.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.bxo_diagnostics.argtypes = [vp, dp]
        L.bxo_run.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp, ctypes.c_int]
        L.bxo_run.restype = ctypes.c_int
        L.bxo_slope.argtypes = [ctypes.c_double] * 3
        L.bxo_slope.restype = ctypes.c_double
        L.bxo_face.argtypes = [ctypes.c_int, dp, dp, dp]
        L.bxo_set_threads.argtypes = [ctypes.c_int]
        L.bxo_get_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleGrid:
    """The reference's global state for one grid, stepped by the C restatement."""

    def __init__(self, prm, ic: Optional[Dict[str, np.ndarray]] = None):
        L = lib()
        self.prm = prm
        P = _Params()
        P.nx, P.ny, P.nz = prm.nx, prm.ny, prm.nz
        P.dx, P.dy, P.dz = prm.dx, prm.dy, prm.dz
        P.gammafix = prm.gammafix
        P.small_rho, P.small_eps, P.vlimit, P.cfl = prm.small_rho, prm.small_eps, prm.vlimit, prm.cfl
        P.xabun, P.yabun, P.zabun, P.mu_z = prm.xabun, prm.yabun, prm.zabun, prm.mu_z
        for i, v in enumerate((prm.avmagx, prm.avmagy, prm.avmagz)):
            P.avmag[i] = v
        for i, v in enumerate((prm.no_outflow_xl, prm.no_outflow_xr, prm.no_outflow_yl,
                               prm.no_outflow_yr, prm.no_outflow_zl, prm.no_outflow_zr)):
            P.no_outflow[i] = int(v)
        self._h = L.bxo_create(ctypes.byref(P))
        n = prm.ngrid
        self.n = n
        self.cons = np.zeros((n, 5))
        self.cons_old = np.zeros((n, 5))
        self.u = np.zeros((n, 3))
        self.p = np.zeros(n)
        self.adindx = np.zeros(n)
        self.qq = np.zeros((n, 3))
        L.bxo_bind(self._h, _dp(self.cons), _dp(self.cons_old), _dp(self.u), _dp(self.p),
                   _dp(self.adindx), _dp(self.qq))
        if ic is not None:
            self.load(ic)

    def load(self, ic: Dict[str, np.ndarray]) -> None:
        self.cons[...] = ic["cons"]
        self.u[...] = ic["u"]
        self.p[...] = ic["p"]
        self.adindx[...] = ic["adindx"]
        self.cons_old[...] = ic.get("cons_old", ic["cons"])

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().bxo_destroy(h)
            self._h = None

    # reference subroutines
    def state(self): lib().bxo_state(self._h)
    def velocity(self): lib().bxo_velocity(self._h)
    def courant(self) -> float: return lib().bxo_courant(self._h)
    def source(self, dt: float): lib().bxo_source(self._h, dt)
    def set_ghost_cells(self): lib().bxo_set_ghost_cells(self._h)
    def flux(self, avg: int, dt: float): lib().bxo_flux(self._h, avg, dt)
    def copy_old(self): lib().bxo_copy_old(self._h)
    def avisc(self): lib().bxo_avisc(self._h)

    def diagnostics(self) -> np.ndarray:
        out = np.zeros(12)
        lib().bxo_diagnostics(self._h, _dp(out))
        return out

    def geometry(self):
        b = np.zeros(self.n, np.int32)
        nb = np.zeros((self.n, 6), np.int32)
        lib().bxo_geometry(self._h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return b, nb

    def run(self, maxsteps: int, starttime: float = 0.0, endtime: Optional[float] = None, diag: int = 0):
        if endtime is None:
            endtime = self.prm.endtime
        dts = np.zeros(max(maxsteps, 1))
        times = np.zeros(max(maxsteps, 1))
        k = lib().bxo_run(self._h, maxsteps, starttime, endtime, _dp(dts), _dp(times), diag)
        return dts[:k].copy(), times[:k].copy()


def slope(c0: float, c1: float, c2: float) -> float:
    return lib().bxo_slope(c0, c1, c2)


def face(idir: int, L, R) -> np.ndarray:
    Lv = np.ascontiguousarray(L, np.float64)
    Rv = np.ascontiguousarray(R, np.float64)
    F = np.zeros(5)
    lib().bxo_face(idir, _dp(Lv), _dp(Rv), _dp(F))
    return F


def set_threads(n: int) -> None:
    lib().bxo_set_threads(n)


def get_threads() -> int:
    return lib().bxo_get_threads()


# ---------------------------------------------------------------- reference driver
def have_reference() -> bool:
    return os.path.exists(REF_HARNESS)


def run_reference(prm, ic: Dict[str, np.ndarray], maxsteps: int, mode: int = 0,
                  threads: int = 4, diag: int = 0, workdir: Optional[str] = None,
                  celldump: int = 0, grav: int = 0, env_extra: Optional[Dict[str, str]] = None
                  ) -> Dict[str, np.ndarray]:
    """Run the reference's own subroutines (oracle/_ref/ref_harness) on ``ic``.

    Returns dt/time per step and the final cons/u/p; with mode=1 also the
    snapshots after every sub-call of step 1; with celldump=1 (binary) or 2
    (ASCII) also the bytes of the celldump that write_files writes after the
    run (``res["celldump"]``); grav=1 (a fixed potential ic["phi"], calc_gforce)
    or 2 (and external_grav) turn on the gravity calls of the main loop;
    env_extra adds environment variables (OpenMP binding) for the run."""
    if not have_reference():
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    from boxzyhydro_amd.ics import write_ic_file

    n = prm.ngrid
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        with open(os.path.join(td, "run.nml"), "w") as f:
            f.write(prm.to_namelist())
        write_ic_file(os.path.join(td, "ic.bin"), ic)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        env.update(env_extra or {})
        if grav and "phi" not in ic:
            raise ValueError("grav needs ic['phi']")
        subprocess.run([REF_HARNESS, "run.nml", "ic.bin", "out", str(maxsteps), str(mode), str(diag),
                        str(celldump), str(grav)], cwd=td, env=env, check=True, stdout=subprocess.DEVNULL)
        dt = np.fromfile(os.path.join(td, "out.dt"), np.float64).reshape(-1, 2)
        fin = np.fromfile(os.path.join(td, "out.final"), np.float64)
        tim = open(os.path.join(td, "out.timing")).read().split()
        res = {"dt": dt[:, 0].copy(), "time": dt[:, 1].copy(),
               "wall_s": float(tim[0]), "threads": int(tim[2]),
               "cons": fin[:5 * n].reshape(n, 5).copy(),
               "u": fin[5 * n:8 * n].reshape(n, 3).copy(),
               "p": fin[8 * n:9 * n].copy()}
        es = np.fromfile(os.path.join(td, "out.eos_state"), np.float64)
        res["adindx"], res["muc"] = es[:n].copy(), es[n:2 * n].copy()
        if mode == 1:
            raw = open(os.path.join(td, "out.snap"), "rb").read()
            rec = 4 + 8 + 8 * (5 * n + 3 * n + n + 5 * n)
            snaps = []
            for k in range(len(raw) // rec):
                b = raw[k * rec:(k + 1) * rec]
                tag = int(np.frombuffer(b[:4], np.int32)[0])
                arr = np.frombuffer(b[4:], np.float64)
                snaps.append({"tag": tag, "dt": float(arr[0]),
                              "cons": arr[1:1 + 5 * n].reshape(n, 5).copy(),
                              "u": arr[1 + 5 * n:1 + 8 * n].reshape(n, 3).copy(),
                              "p": arr[1 + 8 * n:1 + 9 * n].copy(),
                              "cons_old": arr[1 + 9 * n:1 + 14 * n].reshape(n, 5).copy()})
            res["snaps"] = snaps
        if celldump:
            with open(os.path.join(td, f"celldump.{len(dt):08d}"), "rb") as f:
                res["celldump"] = f.read()
        return res


def reference_eos_tables(prm, workdir: Optional[str] = None) -> Dict[str, np.ndarray]:
    """The tables the reference's initialize_eos() + calc_eos_table()
    (eos.f90:30-51, 127-520) leave for ``prm`` (ref_harness mode 3), in the
    reference's units after the call: rho_table in code units, eng_table and
    eng_table2 in code units, the rest as computed.  2-D tables have shape
    (NEOS_RHO, NEOS_T) (C order of the Fortran (NEOS_T, NEOS_RHO) arrays)."""
    if not have_reference():
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        with open(os.path.join(td, "run.nml"), "w") as f:
            f.write(prm.to_namelist())
        subprocess.run([REF_HARNESS, "run.nml", "-", "out", "0", "3"], cwd=td, check=True,
                       stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
        raw = open(os.path.join(td, "out.eos"), "rb").read()
    nt, nr = (int(v) for v in np.frombuffer(raw[:8], np.int32))
    arr = np.frombuffer(raw[8:], np.float64)
    out = {"NEOS_T": nt, "NEOS_RHO": nr, "log_rho_eos_low": float(arr[0]), "drho_eos": float(arr[1])}
    k = 2
    for name, size in (("rho_table", nr), ("deng_eos_array", nr), ("tk_table", nt)):
        out[name] = arr[k:k + size].copy()
        k += size
    for name in ("eng_table", "gamma_table", "muc_table", "eng_table2", "tk_table2", "muc_table2",
                 "gamma_table2"):
        out[name] = arr[k:k + nt * nr].reshape(nr, nt).copy()
        k += nt * nr
    assert k == arr.size
    return out


def reference_state(prm, ic: Dict[str, np.ndarray], workdir: Optional[str] = None) -> Dict[str, np.ndarray]:
    """One call of the reference's state() (state.f90:6-89) on ``ic``
    (ref_harness mode 4): cons, u, p, adindx and muc_array after it."""
    if not have_reference():
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    from boxzyhydro_amd.ics import write_ic_file

    n = prm.ngrid
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        with open(os.path.join(td, "run.nml"), "w") as f:
            f.write(prm.to_namelist())
        write_ic_file(os.path.join(td, "ic.bin"), ic)
        subprocess.run([REF_HARNESS, "run.nml", "ic.bin", "out", "0", "4"], cwd=td, check=True,
                       stdout=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS="1"))
        fin = np.fromfile(os.path.join(td, "out.final"), np.float64)
        es = np.fromfile(os.path.join(td, "out.eos_state"), np.float64)
    return {"cons": fin[:5 * n].reshape(n, 5).copy(), "u": fin[5 * n:8 * n].reshape(n, 3).copy(),
            "p": fin[8 * n:9 * n].copy(), "adindx": es[:n].copy(), "muc": es[n:2 * n].copy()}


def reference_restart(prm, celldump: bytes, step: int, threads: int = 1,
                      workdir: Optional[str] = None) -> Dict[str, np.ndarray]:
    """The reference's read_hydro() (through oracle/_ref/ref_harness, mode 2)
    on the binary celldump ``celldump`` saved as celldump.<step>: the
    cons/u/p it leaves after its state() call."""
    if not have_reference():
        raise RuntimeError("reference driver not built (make -C oracle ref)")
    n = prm.ngrid
    with tempfile.TemporaryDirectory(dir=workdir) as td:
        with open(os.path.join(td, "run.nml"), "w") as f:
            f.write(prm.to_namelist())
        with open(os.path.join(td, f"celldump.{step:08d}"), "wb") as f:
            f.write(celldump)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads))
        subprocess.run([REF_HARNESS, "run.nml", "none", "out", str(step), "2"], cwd=td, env=env, check=True,
                       stdout=subprocess.DEVNULL)
        fin = np.fromfile(os.path.join(td, "out.final"), np.float64)
        return {"cons": fin[:5 * n].reshape(n, 5).copy(), "u": fin[5 * n:8 * n].reshape(n, 3).copy(),
                "p": fin[8 * n:9 * n].copy()}
