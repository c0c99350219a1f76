"""Moving bodies over the device case: rigid motion, the refinement following
the body, and free motion under the fluid force.

The reference sketches a moving-solid path (C_FSI_INTERFACE, Infor_IB and
map_node_IB in header/Grid_Class.h:101-109, 250-253; the node add/remove of
Grid_Manager_dynamic.cpp as its solid points move) without a flow solver around
it.  Here a body is a set of Lagrangian markers moved on the device every coarse
step (mt_markers_rigid), and, once it has left the refinement built around it by
more than `remesh_dist` level-0 cells, a new case is built around its new
position (level sets on the device, mt_levelset_build) and takes the old state
(mt_case_remap).

Units: positions of the body centre in level-0 cells; velocities physical, in
lattice units (the same at every level under the acoustic scaling of the
refinement: a coarse step moves the body by V level-0 cells); the fluid force
on the body is the mean per finest-level collision reported by mt_ib_force, in
finest-level lattice units, and the momentum it gives the body per coarse step,
in level-0 units, is that force x 2^Lmax collisions / 8^Lmax (the mass of a
finest cell is 8^-Lmax of a level-0 cell's).
"""
from __future__ import annotations

import numpy as np

from . import amr
from . import cases as C


class RigidBody:
    """Centre, velocity and angular velocity of a rigid body.  mass=None: the
    motion is prescribed (velocity and angular velocity stay as set); otherwise
    the velocity follows m dV/dt = F + m g, integrated once per coarse step
    (explicit Euler on V, the position then moved with the new V)."""

    def __init__(self, center, velocity=(0.0, 0.0, 0.0), omega=(0.0, 0.0, 0.0), mass=None,
                 gravity=(0.0, 0.0, 0.0)):
        self.center = np.asarray(center, np.float64).copy()
        self.velocity = np.asarray(velocity, np.float64).copy()
        self.omega = np.asarray(omega, np.float64).copy()
        self.mass = mass
        self.gravity = np.asarray(gravity, np.float64).copy()

    def advance(self, impulse, dt=1.0):
        """One step of dt coarse steps: impulse = momentum the fluid gave the body
        (level-0 units) over the step."""
        if self.mass is not None:
            self.velocity = self.velocity + np.asarray(impulse, np.float64) / self.mass + dt * self.gravity
        self.center = self.center + dt * self.velocity


class MarkerBody:
    """A marker body carried through a refined case.

    build(center) returns the case (a dict of cases.*, with markers) for the body
    centred at `center` (level-0 cells); every coarse step the markers move
    rigidly with the body on the device, and when the centre is more than
    remesh_dist level-0 cells from the centre the refinement was built for, the
    case is rebuilt there and takes the state of the old one (cells uncovered
    start at feq(rho0, body velocity)).  forces: per coarse step, the mean fluid
    force per finest collision (physical frame)."""

    def __init__(self, build, body, remesh_dist=0.5, precision=amr.FP64):
        self.build, self.body, self.remesh_dist, self.precision = build, body, remesh_dist, precision
        self.case = build(body.center)
        self.g = amr.AmrCase(self.case, precision=precision)
        self.Lmax = self.case["Lmax"]
        self.mesh_center = body.center.copy()
        x = self.case["markers"]["xyz"]
        self.g.move_markers(x, self._velocities(x))
        self.forces, self.remeshes = [], []

    def _velocities(self, x):
        s = float(1 << self.Lmax)
        r = x - self.body.center * s
        return C.reported(self.body.velocity + np.cross(self.body.omega, r))

    def step(self, n=1):
        s = float(1 << self.Lmax)
        for _ in range(n):
            self.g.step(1)
            f, calls = self.g.ib_force(reset=True)
            mean = f / max(calls, 1)
            self.forces.append(mean)
            old = self.body.center.copy()
            self.body.advance(mean * s / s ** 3)
            self.g.move_markers_rigid(s, old * s, self.body.velocity, self.body.omega)
            if np.max(np.abs(self.body.center - self.mesh_center)) > self.remesh_dist:
                self.remesh()

    def remesh(self):
        x, v = self.g.markers()
        case = self.build(self.body.center)
        h = amr.AmrCase(case, precision=self.precision)
        counts = h.remap_from(self.g, u_fill=C.reported(self.body.velocity))
        h.move_markers(x, v)
        self.g.close()
        self.g, self.case = h, case
        self.mesh_center = self.body.center.copy()
        self.remeshes.append(counts)

    def close(self):
        self.g.close()
