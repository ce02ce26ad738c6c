"""Host driver of the spe10 fork's IMPES loop (``src/spe10.c:97-322``) over
the GPU library: the reservoir object (permeability, porosity, saturation,
pressure, all device-resident) and a preconditioner context with the fork's
settings.  ``Reservoir.run`` is the reference's main loop; the pieces
(``assemble``, ``solve``, ``transport``) are exposed for tests.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pc as P
from ._lib import STEP_HOOK, Spe10Log, check, lib

# Stage tables of main (src/spe10.c:117-119).
SIMU_TIME_STAGES = (0.0, 99.99, 199.99, 399.99, 899.99, 1599.99, 1999.99)
DS_MAX_STAGES = (0.0, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001)
S_W_STEPS_STAGES = (0, 50, 50, 50, 50, 50, 50)
H_CELL = (20.0, 10.0, 2.0)   # ft (src/spe10.c:361-363)


def pc_options(sd, en_lv1=8, en_lv2=4, precision=64, extra=""):
    """The fork's preconditioner: DIAG eigen mass (:990-1012), no NIL shift
    of A_cc, ICC(0) level-2 factor (:509-530); per-direction subdomains."""
    sdx, sdy, sdz = sd
    s = (f"-sdx {sdx} -sdy {sdy} -sdz {sdz} -en_lv1 {en_lv1} -en_lv2 {en_lv2} -eg_op 2 "
         f"-no_shift_A_cc -lv2_icc -pc_precision {precision}")
    return (s + " " + extra).strip()


class Reservoir:
    """``SPE10Ctx`` + ``INTERCtx`` + the PCSHELL context on one GPU."""

    def __init__(self, mesh, perm, poro, sd, H=H_CELL, en_lv1=8, en_lv2=4, precision=64,
                 extra=""):
        self.mesh = tuple(int(v) for v in mesh)
        self.H = tuple(float(h) for h in H)
        M, N, Pz = self.mesh
        perm = np.asarray(perm, dtype=np.float64).reshape(3, -1)
        self.handle = C.c_void_p()
        check(lib().pmx_spe10_create(C.byref(self.handle), np.array(self.mesh, dtype=np.int32),
                                     np.array(self.H), np.ascontiguousarray(perm[0]),
                                     np.ascontiguousarray(perm[1]), np.ascontiguousarray(perm[2]),
                                     np.ascontiguousarray(np.ravel(poro), dtype=np.float64)))
        self.ctx = P.PCCtx()
        dom = (M * self.H[0], N * self.H[1], Pz * self.H[2])
        P.PC_init(self.ctx, dom, self.mesh, pc_options(sd, en_lv1, en_lv2, precision, extra))

    @property
    def ncell(self):
        return int(np.prod(self.mesh))

    def assemble(self):
        check(lib().pmx_spe10_assemble(self.handle, self.ctx.handle))

    def system(self):
        """(kappa (3, n), wd, rhs) of the current saturation, host copies."""
        n = self.ncell
        out = [np.zeros(n) for _ in range(5)]
        check(lib().pmx_spe10_get_system(self.handle, *[o.ctypes.data_as(C.c_void_p) for o in out]))
        return np.stack(out[:3]), out[3], out[4]

    def solve(self, rtol=1e-8, max_it=10000):
        its, rn = C.c_int(0), C.c_double(0.0)
        rc = lib().pmx_spe10_solve(self.handle, self.ctx.handle, rtol, max_it, C.byref(its),
                                   C.byref(rn))
        if rc not in (0, 91):
            check(rc)
        return its.value, rn.value

    def transport(self, steps=50, DS_max=0.001):
        dt = C.c_double(0.0)
        check(lib().pmx_spe10_transport(self.handle, int(steps), float(DS_max), C.byref(dt)))
        return dt.value

    def saturation_range(self):
        lo, hi = C.c_double(0.0), C.c_double(0.0)
        check(lib().pmx_spe10_saturation_range(self.handle, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def state(self):
        sw, p = np.zeros(self.ncell), np.zeros(self.ncell)
        check(lib().pmx_spe10_get_state(self.handle, sw.ctypes.data_as(C.c_void_p),
                                        p.ctypes.data_as(C.c_void_p)))
        return sw, p

    def set_state(self, S_w=None, P_=None):
        a = None if S_w is None else np.ascontiguousarray(S_w, dtype=np.float64).ravel()
        b = None if P_ is None else np.ascontiguousarray(P_, dtype=np.float64).ravel()
        check(lib().pmx_spe10_set_state(self.handle,
                                        None if a is None else a.ctypes.data_as(C.c_void_p),
                                        None if b is None else b.ctypes.data_as(C.c_void_p)))

    def run(self, max_steps=20, rtol=1e-8, max_it=10000, simu_time_stages=SIMU_TIME_STAGES,
            DS_max_stages=DS_MAX_STAGES, steps_stages=S_W_STEPS_STAGES, output_dir=None,
            save_files_freq=1, output_P=False, output_realtime=False):
        """The IMPES loop of main (src/spe10.c:197-274); returns the per-step
        log rows as dicts.  With ``output_dir`` the driver's files are
        written there as the loop runs (``P_solver.csv``, ``S_w<step>.vtr``
        every ``save_files_freq`` steps, ``P<step>.vtr`` with ``output_P``,
        the ``.vtr.series`` indexes); ``output_realtime`` prints each step's
        line as the reference does (:251-253)."""
        from . import io_vtk

        logs = (Spe10Log * max(1, max_steps))()
        done = C.c_int(0)
        out = None
        hook = None
        if output_dir is not None or output_realtime:
            if output_dir is not None:
                out = io_vtk.ImpesOutput(output_dir, self.mesh, self.H, save_files_freq, output_P)
            errors = []

            def on_step(_user, row_p):
                try:
                    row = row_p.contents.as_dict()
                    if output_realtime:
                        wall = row["t_assemble"] + row["t_setup"] + row["t_solve"]
                        print(f"{row['step']}, simu_time={row['simu_time']:.5E}, "
                              f"stage_ind={row['stage']}, DS_max={row['DS_max']:.5f}, "
                              f"S_w_steps={row['S_w_steps']}, wall_time={wall:.5f}, "
                              f"iter_count={row['its']}, S_w=[{row['S_w_min']:.5E}, "
                              f"{row['S_w_max']:.5E}].", flush=True)
                    if out is not None:
                        sw, p = self.state()
                        out.step(row, sw, p)
                    return 0
                except Exception as e:  # surfaced after the loop
                    errors.append(e)
                    return 1

            hook = STEP_HOOK(on_step)
            check(lib().pmx_spe10_set_step_hook(self.handle, C.cast(hook, C.c_void_p), None))
        try:
            rc = lib().pmx_spe10_run(self.handle, self.ctx.handle, int(max_steps),
                                     np.array(simu_time_stages, dtype=np.float64),
                                     np.array(DS_max_stages, dtype=np.float64),
                                     np.array(steps_stages, dtype=np.int32), rtol, max_it,
                                     C.cast(logs, C.c_void_p), C.byref(done))
        finally:
            if hook is not None:
                lib().pmx_spe10_set_step_hook(self.handle, None, None)
            if out is not None:
                out.close()
        if hook is not None and errors:
            raise errors[0]
        rows = [logs[i].as_dict() for i in range(done.value)]
        check(rc)
        return rows

    def close(self):
        if self.handle:
            lib().pmx_spe10_destroy(self.handle)
            self.handle = C.c_void_p()
        if self.ctx.handle is not None:
            P.PC_final(self.ctx)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
