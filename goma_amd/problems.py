"""Problem descriptions for the steady Navier-Stokes (+energy) workloads.

A Problem carries what Goma gets from its input deck and material file for
this path: the mesh, material constants (density, viscosity model, heat
capacity, conductivity, Boussinesq gravity and expansion coefficient) and the
Dirichlet conditions plus the pressure datum.  Variable types use Goma's
numbering (include/rf_fem_const.h:174-200): U1=0, U2=1, U3=2, T=3, P=9.

Nondimensionalisation (Goma has no Ra/Re cards; buoyancy is
rho*g*(-beta*(T-Tref)), src/mm_std_models.c:256-284):
  * lid-driven cavity:        rho = Re, mu = 1, lid speed 1           -> Re
  * natural convection:       rho = 1, Cp = 1, k = 1, mu = Pr,
                              g = (0, -Ra*Pr), beta = 1, dT = 1       -> Ra, Pr
                              (velocities in units of alpha/L)
  * lid cavity with energy:   rho = 1, mu = 1/Re, Cp = 1, k = 1/(Re*Pr),
                              g = (0, -Ra/(Re^2 Pr), 0), beta = 1     -> Re, Ra, Pr
  * Carreau-Yasuda convection: rho = 1, Cp = 1, k = 1, mu0 = 1 (Pr = 1),
                              g = (0, -Ra), beta = 1
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace

import numpy as np

from .mesh import Mesh, structured_mesh

U1, U2, U3, TEMP, PRES = 0, 1, 2, 3, 9


@dataclass
class Physics:
    energy: int = 0
    visc_model: int = 0  # 0 Newtonian, 1 Carreau-Yasuda
    rho: float = 1.0
    mu0: float = 1.0
    muinf: float = 0.0
    lam: float = 0.0
    nexp: float = 1.0
    aexp: float = 2.0
    Cp: float = 1.0
    kcond: float = 1.0
    hsrc: float = 0.0
    g: tuple = (0.0, 0.0, 0.0)
    beta: float = 0.0
    Tref: float = 0.0
    pspg: int = 0          # Pressure Stabilization: 0 no, 1 global
    ps_scaling: float = 0.0  # Pressure Stabilization Scaling
    energy_supg: float = 0.0  # energy weight function SUPG coefficient, 0: Galerkin


@dataclass
class Problem:
    name: str
    mesh: Mesh
    phys: Physics
    dbc_node: np.ndarray  # int32
    dbc_var: np.ndarray  # int32
    dbc_value: np.ndarray  # float64
    datum_elem: int = 0
    datum_value: float = 0.0
    meta: dict = field(default_factory=dict)

    @property
    def dim(self) -> int:
        return self.mesh.dim


class _BC:
    def __init__(self):
        self.n, self.v, self.x = [], [], []

    def add(self, nodes, var, value):
        nodes = np.asarray(nodes, dtype=np.int64)
        self.n.append(nodes)
        self.v.append(np.full(nodes.shape, var, dtype=np.int64))
        self.x.append(np.broadcast_to(np.asarray(value, dtype=np.float64), nodes.shape).copy())

    def arrays(self):
        if not self.n:
            return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0))
        return (
            np.concatenate(self.n).astype(np.int32),
            np.concatenate(self.v).astype(np.int32),
            np.concatenate(self.x).astype(np.float64),
        )


def _walls(mesh: Mesh):
    lat = mesh.lattice()
    m = 2 * mesh.n
    faces = {}
    names = "xyz"
    for k in range(mesh.dim):
        faces[names[k] + "0"] = np.nonzero(lat[:, k] == 0)[0]
        faces[names[k] + "1"] = np.nonzero(lat[:, k] == m)[0]
    boundary = np.nonzero(np.any((lat == 0) | (lat == m), axis=1))[0]
    return lat, faces, boundary


def lid_cavity(n: int = 32, Re: float = 100.0, dim: int = 2, energy: bool = False,
               Ra: float = 0.0, Pr: float = 0.71, jitter: float = 0.0) -> Problem:
    """Lid-driven cavity: no-slip walls, lid y=1 moving with u=1.

    Nodes on the lid's rim (where it meets another wall) keep u=0.  With
    ``energy`` the lid is held at T=+0.5 and the bottom y=0 at T=-0.5, other
    walls adiabatic, with Boussinesq buoyancy along -y.
    """
    mesh = structured_mesh(dim, n, jitter=jitter)
    lat, faces, boundary = _walls(mesh)
    m = 2 * n
    bc = _BC()
    for a in range(dim):
        bc.add(boundary, a, 0.0)
    lid = faces["y1"]
    inner = np.ones(lid.shape, dtype=bool)
    for k in range(dim):
        if k == 1:
            continue
        inner &= (lat[lid, k] > 0) & (lat[lid, k] < m)
    bc.add(lid[inner], U1, 1.0)
    if energy:
        phys = Physics(energy=1, rho=1.0, mu0=1.0 / Re, Cp=1.0, kcond=1.0 / (Re * Pr),
                       g=(0.0, -Ra / (Re * Re * Pr), 0.0), beta=1.0, Tref=0.0)
        bc.add(faces["y1"], TEMP, 0.5)
        bc.add(faces["y0"], TEMP, -0.5)
    else:
        phys = Physics(energy=0, rho=Re, mu0=1.0)
    nn, vv, xx = bc.arrays()
    return Problem(name=f"lid_cavity_{dim}d_n{n}", mesh=mesh, phys=phys, dbc_node=nn, dbc_var=vv,
                   dbc_value=xx, datum_elem=0, datum_value=0.0,
                   meta=dict(Re=Re, Ra=Ra, Pr=Pr, energy=energy))


def natural_convection(n: int = 256, Ra: float = 1e5, Pr: float = 0.71, dim: int = 2,
                       carreau: bool = False, jitter: float = 0.0, seed: int = 20261202) -> Problem:
    """Differentially heated cavity: T=+0.5 at x=0, T=-0.5 at x=1, other walls
    adiabatic, no-slip everywhere, gravity along -y.

    ``carreau`` switches to the shear-thinning Carreau-Yasuda fluid
    (eta0=1, eta_inf=0.01, lambda=1, n=0.5, a=2) with Pr = 1 in eta0 units.
    """
    mesh = structured_mesh(dim, n, jitter=jitter, seed=seed)
    lat, faces, boundary = _walls(mesh)
    bc = _BC()
    for a in range(dim):
        bc.add(boundary, a, 0.0)
    bc.add(faces["x0"], TEMP, 0.5)
    bc.add(faces["x1"], TEMP, -0.5)
    if carreau:
        phys = Physics(energy=1, visc_model=1, rho=1.0, mu0=1.0, muinf=0.01, lam=1.0, nexp=0.5,
                       aexp=2.0, Cp=1.0, kcond=1.0, g=(0.0, -Ra, 0.0), beta=1.0, Tref=0.0)
    else:
        phys = Physics(energy=1, rho=1.0, mu0=Pr, Cp=1.0, kcond=1.0, g=(0.0, -Ra * Pr, 0.0),
                       beta=1.0, Tref=0.0)
    nn, vv, xx = bc.arrays()
    return Problem(name=f"natconv_{dim}d_n{n}", mesh=mesh, phys=phys, dbc_node=nn, dbc_var=vv,
                   dbc_value=xx, datum_elem=0, datum_value=0.0,
                   meta=dict(Ra=Ra, Pr=Pr, carreau=carreau, jitter=jitter, seed=seed))


def config(idx: int, n: int | None = None, Ra: float | None = None) -> Problem:
    """The five workloads, by index 0..4 (sizes overridable for tests)."""
    if idx == 0:
        return lid_cavity(n=n or 32, Re=100.0)
    if idx == 1:
        return natural_convection(n=n or 256, Ra=Ra or 1e5, Pr=0.71)
    if idx == 2:
        return lid_cavity(n=n or 48, Re=400.0, dim=3, energy=True, Ra=Ra or 1e4, Pr=0.71)
    if idx in (3, 4):
        return natural_convection(n=n or 128, Ra=Ra or 1e3, carreau=True, jitter=0.15)
    raise ValueError(idx)


def physics_at(prob: Problem, Ra: float) -> Physics:
    """The problem's material constants with the Rayleigh number changed (the
    continuation parameter; Goma changes the gravity component of the momentum
    source through update_parameterC, src/ac_update_parameter.c:48)."""
    m = prob.meta
    if "Ra" not in m or not prob.phys.energy:
        raise ValueError(f"{prob.name} has no Rayleigh number")
    if "Re" in m:
        g = -Ra / (m["Re"] ** 2 * m["Pr"])
    elif m.get("carreau"):
        g = -Ra
    else:
        g = -Ra * m["Pr"]
    return replace(prob.phys, g=(0.0, g, 0.0))


def sweep_ra(count: int = 64, lo: float = 1e3, hi: float = 1e6) -> np.ndarray:
    """Log-spaced continuation parameter values."""
    return np.logspace(np.log10(lo), np.log10(hi), count)
