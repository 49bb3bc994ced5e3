"""Run parameters of the hydro step, mirroring the reference namelist groups.

The reference reads six mandatory namelist groups from ``argv[1]``
(input.f90:7-16, read in input.f90:21-36).  Only the keys that reach the hydro
step are carried here; ``to_namelist`` writes a complete file that the
reference's ``read_params`` accepts (all six groups present), and
``from_namelist`` reads one back.

Defaults are the reference's (parameters.f90:26-93) except where a field is
documented otherwise.
"""
from __future__ import annotations

import dataclasses
import re
from typing import Dict, Tuple


@dataclasses.dataclass
class HydroParams:
    # &grid_params (input.f90:9-11); nx, ny, nz INCLUDE two ghost layers per side
    nx: int = 50
    ny: int = 50
    nz: int = 50
    dx: float = 0.25
    dy: float = 0.25
    dz: float = 0.25
    no_outflow_xl: bool = True
    no_outflow_xr: bool = True
    no_outflow_yl: bool = True
    no_outflow_yr: bool = True
    no_outflow_zl: bool = True
    no_outflow_zr: bool = True
    # flux_ang (flux_ang.f90) instead of flux, about the z axis through y = -yoffset
    fluxangmom: bool = False
    yoffset: float = 0.0
    # &hydro (input.f90:12-13)
    small_rho: float = 1e-10
    small_eps: float = 1e-40
    vlimit: float = 10.0
    endtime: float = 0.0
    starttime: float = 0.0
    dtout: float = 10.0
    cfl: float = 0.1
    avmagx: float = 0.0
    avmagy: float = 0.0
    avmagz: float = 0.0
    # &eos_input (input.f90:14-15).  H2STAT=-1: fixed gamma; 0: H2 with a fixed
    # ortho:para ratio ac:bc; 1: H2 in ortho/para equilibrium (eos.f90:179-290).
    # The tabulated forms (0, 1) also use the table keys below.
    gammafix: float = 1.66666666667
    xabun: float = 0.73
    yabun: float = 0.27
    zabun: float = 0.0
    mu_z: float = 16.78
    H2STAT: int = -1
    NEOS_RHO: int = 1000
    dTk_eos: float = 5.0
    tk_bgrnd: float = 1e-6
    brot: float = 85.4
    vib: float = 5987.0
    ac: float = 1.0
    bc: float = 3.0

    @property
    def ngrid(self) -> int:
        return self.nx * self.ny * self.nz

    @property
    def shape(self) -> Tuple[int, int, int]:
        """(nz, ny, nx): C-order shape of one field plane (x fastest)."""
        return (self.nz, self.ny, self.nx)

    @property
    def no_outflow(self) -> Tuple[bool, ...]:
        z_ok = self.nz >= 6  # init_grid.f90:76-79 drops z walls on a flat z axis
        return (self.no_outflow_xl, self.no_outflow_xr, self.no_outflow_yl,
                self.no_outflow_yr, self.no_outflow_zl and z_ok,
                self.no_outflow_zr and z_ok)

    @property
    def interior_cells(self) -> int:
        ix = self.nx - 4
        iy = self.ny - 4
        iz = self.nz - 4 if self.nz >= 6 else self.nz
        return ix * iy * iz

    @property
    def muc(self) -> float:
        """Mean molecular weight, eos.f90:144."""
        return 1.0 / (self.xabun * 0.5 + self.yabun * 0.25 + self.zabun / self.mu_z)

    @property
    def rgas(self) -> float:
        """Gas constant in code units, units.f90:10-17."""
        pi = 3.1415926535897931
        time = (3600.0 * 365.25 * 24.0) / (2.0 * pi)
        vel = 1.496e13 / time
        return 8.254e7 / ((vel * vel) / 1.0)

    def to_namelist(self) -> str:
        def f(x: float) -> str:
            return repr(float(x)).replace("e", "d") if "e" in repr(float(x)) else repr(float(x)) + "d0"

        def b(x: bool) -> str:
            return ".true." if x else ".false."

        return "\n".join([
            "&io_control", "write_ascii=.false.", "/", "",
            "&grid_params",
            f"nx={self.nx}", f"ny={self.ny}", f"nz={self.nz}",
            f"dx={f(self.dx)}", f"dy={f(self.dy)}", f"dz={f(self.dz)}",
            f"yoffset={f(self.yoffset)}", f"fluxangmom={b(self.fluxangmom)}",
            f"no_outflow_xl={b(self.no_outflow_xl)}", f"no_outflow_xr={b(self.no_outflow_xr)}",
            f"no_outflow_yl={b(self.no_outflow_yl)}", f"no_outflow_yr={b(self.no_outflow_yr)}",
            f"no_outflow_zl={b(self.no_outflow_zl)}", f"no_outflow_zr={b(self.no_outflow_zr)}",
            "/", "",
            "&hydro",
            f"small_rho={f(self.small_rho)}", f"small_eps={f(self.small_eps)}",
            f"vlimit={f(self.vlimit)}", f"endtime={f(self.endtime)}",
            f"starttime={f(self.starttime)}", f"dtout={f(self.dtout)}",
            f"cfl={f(self.cfl)}", "irestart=0",
            f"avmagx={f(self.avmagx)}", f"avmagy={f(self.avmagy)}", f"avmagz={f(self.avmagz)}",
            "/", "",
            "&gravity", "yml_max=10", "/", "",
            "&eos_input",
            f"H2STAT={self.H2STAT}", f"gammafix={f(self.gammafix)}",
            f"xabun={f(self.xabun)}", f"yabun={f(self.yabun)}", f"zabun={f(self.zabun)}",
            f"mu_z={f(self.mu_z)}",
            *self._table_keys(f),
            "/", "",
            "&particle_input", "npart=0", "/", "",
        ])

    def _table_keys(self, f):
        if self.H2STAT < 0:
            # tiny unused tables: H2STAT=-1 only needs muc (eos.f90:144), and
            # main.f90:70-71 builds the tables whatever H2STAT is
            return ["NEOS_RHO=4", "dTk_eos=1d4"]
        return [f"NEOS_RHO={self.NEOS_RHO}", f"dTk_eos={f(self.dTk_eos)}",
                f"tk_bgrnd={f(self.tk_bgrnd)}", f"brot={f(self.brot)}", f"vib={f(self.vib)}",
                f"ac={f(self.ac)}", f"bc={f(self.bc)}"]

    @classmethod
    def from_namelist(cls, text: str) -> "HydroParams":
        """Parse the keys this class carries from Fortran namelist text."""
        fields = {fl.name: fl for fl in dataclasses.fields(cls)}
        lower = {k.lower(): k for k in fields}
        vals: Dict[str, object] = {}
        body = re.sub(r"!.*", "", text)
        for key, val in re.findall(r"([A-Za-z_][A-Za-z0-9_]*)\s*=\s*([^,\n/]+)", body):
            name = lower.get(key.lower())
            if name is None:
                continue
            val = val.strip()
            typ = fields[name].type
            if typ in ("bool", bool):
                vals[name] = val.lower().strip(".").startswith("t")
            elif typ in ("int", int):
                vals[name] = int(val)
            else:
                vals[name] = float(val.lower().replace("d", "e"))
        return cls(**vals)


def sod_params(nx_interior: int = 400, endtime: float = 0.2) -> HydroParams:
    """Sod shock tube: interior nx x 4 x 4 on x in [0,1], gamma=1.4, CFL 0.5, outflow."""
    d = 1.0 / nx_interior
    return HydroParams(nx=nx_interior + 4, ny=8, nz=4, dx=d, dy=d, dz=d,
                       no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False,
                       no_outflow_yr=False, no_outflow_zl=False, no_outflow_zr=False,
                       small_rho=1e-12, small_eps=1e-40, vlimit=10.0, endtime=endtime,
                       dtout=1e30, cfl=0.5, gammafix=1.4)


def box_params(n_interior: int = 256, endtime: float = 1e30) -> HydroParams:
    """3-D box: n^3 interior cells, dx=1, gamma=5/3, CFL 0.5, outflow."""
    return HydroParams(nx=n_interior + 4, ny=n_interior + 4, nz=n_interior + 4,
                       dx=1.0, dy=1.0, dz=1.0,
                       no_outflow_xl=False, no_outflow_xr=False, no_outflow_yl=False,
                       no_outflow_yr=False, no_outflow_zl=False, no_outflow_zr=False,
                       small_rho=1e-10, small_eps=1e-40, vlimit=10.0, endtime=endtime,
                       dtout=1e30, cfl=0.5, gammafix=5.0 / 3.0)
