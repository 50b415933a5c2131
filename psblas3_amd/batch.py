"""Batched mixed-precision solves over the psb_hip_batch_* entry points.

Many independent systems of different sizes advance together on the device:
a step runs check_every FP32 iterations of every live system plus one FP64
transition pass, as one graph launch.  Systems are admitted and retired
between steps; each system's vectors live in the engine's arena, and
``submit`` returns None while no slot or arena room is free (admission by
memory footprint).  The reference has no batched solver; the per-system
iteration is psb_dcg / psb_dcgstab (krylov/psb_dcg.F90:241-294,
krylov/psb_dcgstab.f90:248-364) with the FP64 stop test of istop 2.
"""
import ctypes as C

import numpy as np

from . import _native as N


def _vdev(v):
    """Device handle of a psb_d_vect_type or psb_d_hip_vect (synced to the device)."""
    inner = v.v if hasattr(v, "v") else v
    inner.ensure_dev()
    return inner.dev


class psb_hip_batch:
    """Engine handle.  method "CG" or "BICGSTAB"; refine "reliable" (one FP32
    recurrence with FP64 residual replacement whenever the recursive residual
    drops by ``delta``) or, for BiCGSTAB, "restart" (restarted refinement:
    each cycle solves A d = r/||r|| in FP32 to ``inner_eps``)."""

    def __init__(self, method="CG", max_systems=256, arena_bytes=1 << 30, check_every=8,
                 delta=1e-4, inner_eps=1e-4, use_graph=True, refine="reliable"):
        self.method = method.upper()
        if self.method not in ("CG", "BICGSTAB") or refine not in ("reliable", "restart") or \
                (self.method == "CG" and refine == "restart"):
            raise N.PsbError(N.INVALID, "psb_hip_batch", f"method {method} / refine {refine}")
        code = 0 if self.method == "CG" else (1 if refine == "restart" else 2)
        h = C.c_void_p()
        N.call("psb_hip_batch_create", C.byref(h), code,
               int(max_systems), int(arena_bytes), int(check_every), float(delta),
               float(inner_eps), 1 if use_graph else 0)
        self.h = h
        self.max_systems = max_systems
        self.check_every = check_every
        self._keep = []          # registered matrices / preconditioners stay alive
        self._n = {}             # mat_id -> rows

    # ------------------------------------------------------------ registration
    def register(self, a, prec=None):
        """Register an HLG double matrix (psb_dspmat_type or the HLG object)
        with its DIAG preconditioner (or None).  Returns the matrix id."""
        inner = a.a if hasattr(a, "a") else a
        if inner.get_fmt() != "HLG" or inner._dtype != np.float64:
            raise N.PsbError(N.UNSUPPORTED, "psb_hip_batch.register", "needs a double HLG matrix")
        if prec is not None and prec.ptype not in ("DIAG", "NONE"):
            raise N.PsbError(N.UNSUPPORTED, "psb_hip_batch.register",
                             f"preconditioner {prec.ptype} (DIAG or NONE)")
        diag = None
        if prec is not None and prec.ptype == "DIAG":
            diag = prec.dv.dev
        mid = C.c_int(-1)
        N.call("psb_hip_batch_register", self.h, inner.dev, diag, C.byref(mid))
        inner._has_single = True
        self._keep.append((a, prec))
        self._n[mid.value] = inner.nr
        return mid.value

    def footprint(self, mat_id):
        """Arena bytes one solve on matrix mat_id occupies."""
        return int(N.fn("psb_hip_batch_footprint")(self.h, int(mat_id)))

    # ------------------------------------------------------------ systems
    def can_admit(self, mat_id):
        """True when a solve on mat_id would be admitted now."""
        return bool(N.fn("psb_hip_batch_can_admit")(self.h, int(mat_id)))

    def submit(self, mat_id, b, x0=None, eps=1e-8, itmax=10000):
        """Admit one solve (b and x0 are copied).  Returns its slot id, or None
        when no slot or arena room is free now."""
        sid = C.c_int(-1)
        rc = N.fn("psb_hip_batch_submit")(self.h, int(mat_id), _vdev(b),
                                          _vdev(x0) if x0 is not None else None, float(eps),
                                          int(itmax), C.byref(sid))
        if rc == N.BUSY:
            return None
        if rc != N.SUCCESS:
            detail = N.load().psb_hip_last_error()
            raise N.PsbError(rc, "psb_hip_batch_submit", detail.decode() if detail else "")
        return sid.value

    def step(self):
        N.call("psb_hip_batch_step", self.h)

    def poll(self, wait_last=False):
        """Finished systems as of the step before the last one queued (or the
        last one): list of dicts (id, iterations, outer, relres, status) and
        the number of systems still running."""
        m = self.max_systems
        ids, its, outer, status = (C.c_int * m)(), (C.c_int * m)(), (C.c_int * m)(), (C.c_int * m)()
        rel = (C.c_double * m)()
        nout, nlive = C.c_int(0), C.c_int(0)
        N.call("psb_hip_batch_poll", self.h, 1 if wait_last else 0, m, ids, its, outer, rel,
               status, C.byref(nout), C.byref(nlive))
        done = [{"id": ids[k], "iterations": its[k], "outer": outer[k], "relres": rel[k],
                 "status": status[k]} for k in range(nout.value)]
        return done, nlive.value

    def result(self, sys_id, x=None):
        """Copy a finished system's solution into x (a double vector) and free
        its slot."""
        N.call("psb_hip_batch_result", self.h, int(sys_id), _vdev(x) if x is not None else None)
        if x is not None:
            (x.v if hasattr(x, "v") else x).set_dev()

    def free(self):
        if self.h:
            N.call("psb_hip_batch_free", self.h)
            self.h = None
            self._keep = []

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    # ------------------------------------------------------------ driver
    def solve(self, jobs, eps=1e-8, itmax=10000, on_done=None):
        """Solve jobs = [(mat_id, b, x_out)] with continuous admission: each job
        is submitted as soon as a slot and arena room are free, steps run
        until every job has finished, and each solution is copied into its
        x_out.  Returns per-job dicts (id, iterations, outer, relres, status,
        admitted_at_step)."""
        pending = list(range(len(jobs)))
        slot_job = {}
        out = [None] * len(jobs)
        step = 0
        while pending or slot_job:
            while pending:
                j = pending[0]
                mid, b, _ = jobs[j]
                sid = self.submit(mid, b, eps=eps, itmax=itmax)
                if sid is None:
                    if not slot_job:
                        raise N.PsbError(N.OUTOFMEM, "psb_hip_batch.solve",
                                         f"job {j} does not fit in an empty engine")
                    break
                slot_job[sid] = (j, step)
                pending.pop(0)
            self.step()
            step += 1
            done, _ = self.poll(wait_last=not pending)
            for d in done:
                j, s0 = slot_job.pop(d["id"])
                self.result(d["id"], jobs[j][2])
                d = dict(d, admitted_at_step=s0)
                out[j] = d
                if on_done:
                    on_done(j, d)
        return out
