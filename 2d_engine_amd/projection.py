"""Host-side mirror of the reference's PROJECTION_2D interface, backed by the HIP library.

Reference: PROJECTION_2D.h (class at :8-1536).  Names, fields and argument
meaning follow it so a caller of the reference finds the same surface:

  PROJECTION_2D(levelset, u_mac_x, u_mac_y)      ctor, PROJECTION_2D.h:107-126
  InitializeFromBlock(block)                     :132-232 (tolerance, max_iteration,
                                                 air_bubble_rising / water_drop,
                                                 poisson_solver_type, BC flags)
  fields water_density, air_density, surface_tension, use_delta_function_formulation,
         CSF_model, air_water_simulation          written by the caller,
                                                 EULERIAN_FLUID_SOLVER_2D.h:506-515
  Solve(dt)                                      :235-242
  DetermineDivergence(dt)                        :395-478 -> divergence_field
  DeterminePressure()                            :480-509 (the solve for divergence_field)
  pressure_field, max_velocity_x, max_velocity_y outputs (:29, :71-72); pressure_field in
                                                 the reference's gauge p(0,0) = 0

`tolerance` keeps the reference's absolute rule ||b - A p||_2 < tolerance
(CG_METHOD.cpp:52).  Two further rules, off (0) by default, may be added:
relative_tolerance (||r||_2 <= it * ||b||_2) and div_tolerance (max|D u_proj|).

Error behaviour: like the reference, a solve that hits max_iteration returns
silently (its stats say so); invalid input and device failures raise P2DError.
Arrays are numpy (host) or torch GPU tensors (used in place on the device).
"""
import numpy as np

from . import _native as N

_SOLVER_TYPES = {"CG": N.FP64, "PCG": N.FP64, "MULTIGRID": N.FP64, "MG": N.FP64, "MIXED": N.MIXED}


class PROJECTION_2D:
    def __init__(self, water_levelset, water_velocity_field_mac_x, water_velocity_field_mac_y, h=None):
        phi = water_levelset
        ny, nx = phi.shape
        self.water_levelset = phi
        self.water_velocity_field_mac_x = water_velocity_field_mac_x
        self.water_velocity_field_mac_y = water_velocity_field_mac_y
        self.nx, self.ny = nx, ny
        self.dx = self.dy = (1.0 / nx) if h is None else h
        # defaults of the reference constructor / InitializeFromBlock
        self.tolerance = 1e-4
        self.max_iteration = 30
        self.poisson_solver_type = "PCG"
        self.air_water_simulation = True
        self.air_bubble_rising = True
        self.water_drop = False
        self.Neumann_Boundary_Condition = True
        self.Dirichlet_Boundary_Condition = False
        self.use_delta_function_formulation = True
        self.use_jump_condition_on_viscosity = False
        self.CSF_model = False
        self.water_density = 1000.0
        self.air_density = 1.226
        self.surface_tension = 0.0728
        self.relative_tolerance = 0.0
        self.div_tolerance = 0.0
        self.divergence_field = None
        self.pressure_field = None
        self.max_velocity_x = 0.0
        self.max_velocity_y = 0.0
        self.stats = None
        self._ctx = None
        self._ctx_key = None

    def InitializeFromBlock(self, block):
        """Keys as in the PROJECTION block of 2D_script.sim:311-330 (a dict here)."""
        self.tolerance = float(block.get("tolerance", 1e-4))
        self.max_iteration = int(block.get("max_iteration", 30))
        self.air_bubble_rising = bool(block.get("air_bubble_rising", False))
        self.water_drop = bool(block.get("water_drop", False))
        self.Dirichlet_Boundary_Condition = bool(block.get("Dirichlet_Boundary_Condition", False))
        self.Neumann_Boundary_Condition = bool(block.get("Neumann_Boundary_Condition", False))
        self.poisson_solver_type = str(block.get("poisson_solver_type", "PCG")).upper()
        self.relative_tolerance = float(block.get("relative_tolerance", self.relative_tolerance))
        self.div_tolerance = float(block.get("div_tolerance", self.div_tolerance))

    # ------------------------------------------------------------------ internals
    def _context(self):
        if self.Dirichlet_Boundary_Condition or not self.Neumann_Boundary_Condition:
            raise NotImplementedError("only the no-flux (Neumann) box of PROJECTION_2D.h:1010-1047 is built")
        if self.use_jump_condition_on_viscosity:
            raise NotImplementedError("viscous jump (PROJECTION_2D.h:837-865) is not built")
        prec = _SOLVER_TYPES.get(self.poisson_solver_type)
        if prec is None:
            raise ValueError(f"unknown poisson_solver_type {self.poisson_solver_type}")
        key = (prec, self.tolerance, self.max_iteration, self.relative_tolerance, self.div_tolerance)
        if self._ctx is None or self._ctx_key != key:
            self._ctx = N.Context(self.nx, self.ny, self.dx, precision=prec, abs_tol=self.tolerance,
                                  rtol=self.relative_tolerance, div_tol=self.div_tolerance,
                                  max_iterations=max(1, self.max_iteration))
            self._ctx_key = key
        return self._ctx

    def _densities(self):
        if self.air_bubble_rising:   # PROJECTION_2D.h:330-343
            return self.air_density, self.water_density
        if self.water_drop:          # :346-359
            return self.water_density, self.air_density
        raise ValueError("set air_bubble_rising or water_drop")

    def _sigma(self):
        if self.use_delta_function_formulation and not self.CSF_model:  # :820-834
            return self.surface_tension
        return 0.0

    # ------------------------------------------------------------------ PROJECTION_2D API
    def Solve(self, dt):
        ctx = self._context()
        rho_m, rho_p = self._densities()
        ctx.set_fluid(self.water_levelset, rho_m, rho_p, self._sigma())
        u, v, st = ctx.project(self.water_velocity_field_mac_x, self.water_velocity_field_mac_y, dt,
                               allow_not_converged=True)
        if not hasattr(self.water_velocity_field_mac_x, "data_ptr"):
            self.water_velocity_field_mac_x[...] = u
            self.water_velocity_field_mac_y[...] = v
        self.max_velocity_x, self.max_velocity_y = st["max_u"], st["max_v"]
        self.stats = st
        self.pressure_field = ctx.pressure(mean_zero=False)
        return st

    def DetermineDivergence(self, dt):
        """divergence_field = (D u) / dt of the current face velocities (on the GPU)."""
        ctx = self._context()
        self.divergence_field = ctx.divergence(self.water_velocity_field_mac_x, self.water_velocity_field_mac_y, dt)
        return self.divergence_field

    def DeterminePressure(self):
        """The ghost-fluid pressure solve for whatever divergence_field holds (set by
        DetermineDivergence or by the caller); the velocity fields are not touched."""
        if self.divergence_field is None:
            raise ValueError("divergence_field is not set (call DetermineDivergence or assign it)")
        ctx = self._context()
        rho_m, rho_p = self._densities()
        ctx.set_fluid(self.water_levelset, rho_m, rho_p, self._sigma())
        self.pressure_field, self.stats = ctx.solve_pressure(self.divergence_field, allow_not_converged=True)
        return self.pressure_field
