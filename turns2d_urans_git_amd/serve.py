"""Continuous-batching request loop over libot2d.

Requests (steady polar points and time-spectral pitching cases) arrive over
time.  Between iterations the loop

* admits waiting requests, in arrival order: a request first takes the
  slot of a finished case in a resident cohort of its structure whose
  memory it fits (``ot2d_batch_admit``: the slot is cleared and re-made
  for the new case, the other cases keep running); the rest, while the
  device memory they need (``ot2d_footprint``) fits the budget, become new
  batches (``ot2d_batch_create``), one *cohort* per structure (model,
  precision, instance count) and admission moment, each with its own HIP
  stream;
* advances each resident cohort by ``steps_per_tick`` iterations with
  ``ot2d_iterate_async`` as soon as its previous ones are done
  (``ot2d_query`` never waits), so cohorts advance independently and their
  kernels overlap on the device (the library orders the FP32 LU-SGS work of
  different cohorts among themselves unless ``ot2d_set_fp32_order(0)``
  lifts that; FP64 work overlaps with everything);
* reads a cohort's per-case states (``ot2d_batch_status``) once its steps
  are done: a case whose residual has reached the tolerance stopped on the
  device (status ``CONVERGED``, set by the step itself before its update), and its
  time-to-converge is recorded and its slot becomes free; a cohort with no
  case left running, whose free slots no waiting request took, is destroyed
  and its memory returns to the budget.

State never leaves the device while a case runs.  Cases never interact: a
case gives the same result alone and in any mix (the kernels of one case
read only its own records).

The reference has no request loop; its driver runs one case for ``nsteps``
(src/Turns_overset.f:378-560).  The convergence measure is its residual
norm rsum (src/step.f:60-93).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import json
import math
import time
from collections import deque
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import abi, grid


# ------------------------------------------------------------------ requests
@dataclasses.dataclass
class Request:
    rid: int
    arrival: float                  # seconds after the stream starts
    deck: Dict[str, float]          # ot2d_params fields
    grid: grid.CGrid
    kind: str = "polar"             # "polar" (steady) or "pitch" (time-spectral)

    @property
    def nspec(self) -> int:
        return int(self.deck.get("nspec", 1)) if self.deck.get("timespectral") else 1

    @property
    def cells(self) -> int:
        return (self.grid.jmax - 1) * (self.grid.kmax - 1) * self.nspec

    def key(self):
        d = self.deck
        return (int(d.get("timespectral", 0)), self.nspec, int(d.get("iturb", 1)), int(d.get("lamin", 0)),
                int(d.get("invisc", 0)), int(d.get("lu_fp32", 0)))


@dataclasses.dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float = math.nan
    finished: float = math.nan
    steps: int = 0
    status: int = 0                 # abi.CONVERGED, or the failure code
    cells: int = 0
    forces: Optional[np.ndarray] = None   # (nspec, 6): cl cd cm cpower cfx cfy
    state: Optional[np.ndarray] = None    # final q when kept

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


# ------------------------------------------------------------- device memory
def device_memory():
    """(free, total) bytes of the current HIP device."""
    hip = C.CDLL("libamdhip64.so")
    free, total = C.c_size_t(), C.c_size_t()
    rc = hip.hipMemGetInfo(C.byref(free), C.byref(total))
    if rc != 0:
        raise RuntimeError(f"hipMemGetInfo failed ({rc})")
    return free.value, total.value


# -------------------------------------------------------------------- server
def _params(lib, deck) -> abi.Params:
    p = abi.Params()
    lib.ot2d_default_params(C.byref(p))
    p.update(**deck)
    return p


def _mesh(g) -> abi.MeshDesc:
    return abi.mesh_desc(g.jmax, g.kmax, grid.cgrid_bc(g.jtail1))


class _Cohort:
    """One batch of a structure (model, precision, instance count) on its own
    stream.  Without `slot_grid` its slots are the requests that made it,
    each sized to its request.  With `slot_grid` (the structure's largest
    grid) it has `slots` slots of that size: they start stopped (the grid
    set, then retired) and the requests that made it are admitted into the
    first ones, so later arrivals of any size of the structure find free
    slots in a resident cohort instead of starting a small one."""

    def __init__(self, lib, reqs: List[Request], tol: float, errfn, tune_steps: int = 0,
                 slots: int = 0, slot_grid=None):
        self.lib = lib
        self.tol = tol
        self.key = reqs[0].key()
        if slot_grid is None:
            self.reqs = list(reqs)
            self.case = abi.CaseAPI(lib, "ot2d_", [_params(lib, r.deck) for r in reqs], [_mesh(r.grid) for r in reqs],
                                    errfn=errfn, batch=True)
            for i, r in enumerate(reqs):
                self.case.set_grid(i, r.grid.xv, r.grid.yv)
            self.case.set_convergence(tol)
            self.open = set(range(len(reqs)))
            self.free = set()                      # slots whose case has finished
            # the mesh each slot was allocated for: a request fits a slot when
            # its grid is no larger in either direction
            self.cap = [(r.grid.jmax, r.grid.kmax) for r in reqs]
        else:
            n = max(int(slots), len(reqs))
            self.reqs = [None] * n
            self.case = abi.CaseAPI(lib, "ot2d_", [_params(lib, reqs[0].deck)] * n, [_mesh(slot_grid)] * n,
                                    errfn=errfn, batch=True)
            for i in range(n):
                self.case.set_grid(i, slot_grid.xv, slot_grid.yv)
                self.case.batch_retire(i)
            self.case.set_convergence(tol)
            self.open = set()
            self.free = set(range(n))
            self.cap = [(slot_grid.jmax, slot_grid.kmax)] * n
            for i, r in enumerate(reqs):
                self.admit(i, r)
        self.busy = False                          # steps enqueued and not yet done
        self.dirty = True                          # cases stopped or admitted since the last compaction
        if tune_steps > 0:
            # launch choices of the cohort's shape class (cached per class:
            # only the first cohort of a class measures)
            self.case.autotune(tune_steps)
        self.bytes = self.case.bytes()

    def fits(self, i: int, r: Request) -> bool:
        return r.key() == self.key and r.grid.jmax <= self.cap[i][0] and r.grid.kmax <= self.cap[i][1]

    def admit(self, i: int, r: Request):
        """Put request r into the free slot i (ot2d_batch_admit, then its grid
        and tolerance)."""
        self.case.batch_admit(i, _params(self.lib, r.deck), _mesh(r.grid))
        self.case.set_grid(i, r.grid.xv, r.grid.yv)
        self.case.set_convergence(self.tol, i)
        self.reqs[i] = r
        self.free.discard(i)
        self.open.add(i)
        self.dirty = True

    def close(self):
        self.case.close()


class Server:
    """Continuous-batching loop; see the module docstring.

    ``budget``: device bytes the resident cohorts may hold (default: 80 % of
    the free memory at construction).  ``max_steps``: a case still running
    after that many iterations is retired as not converged.  ``max_resident``:
    at most that many cases resident at once (None: the memory budget alone
    limits them); past it, arrivals wait in arrival order, so an overloaded
    stream queues instead of slowing every resident case.  ``tune_steps``
    > 0: each new cohort takes its launch choices from ``ot2d_autotune``
    (that many steps per candidate, measured once per shape class).
    ``slot_dims``: a function giving a request's structure's largest grid
    (jmax, kmax); with it every new cohort has ``max_cohort`` slots of that
    size (see ``_Cohort``), and a cohort left without running cases stays
    resident while requests are still to arrive."""

    def __init__(self, lib, budget: Optional[int] = None, steps_per_tick: int = 50, tol: float = 1e-10,
                 max_cohort: int = 32, max_steps: int = 100000, keep_state: bool = False, errfn=None,
                 tune_steps: int = 0, max_resident: Optional[int] = None, slot_dims=None):
        self.lib = lib
        abi.bind(lib, "ot2d_")
        lib.ot2d_default_params.argtypes = [C.POINTER(abi.Params)]
        self.budget = int(budget if budget is not None else 0.8 * device_memory()[0])
        self.k = int(steps_per_tick)
        self.tol = float(tol)
        self.max_cohort = int(max_cohort)
        self.max_steps = int(max_steps)
        self.keep_state = keep_state
        self.errfn = errfn
        self.tune_steps = int(tune_steps)
        self.max_resident = max_resident
        self.slot_dims = slot_dims
        self._slot_grids: Dict[tuple, grid.CGrid] = {}
        self.peak_bytes = 0
        self.ticks = 0
        self.slot_admissions = 0   # requests that took a finished case's slot
        self.cohorts_made = 0

    def footprint(self, reqs: Sequence[Request]) -> int:
        params = []
        for r in reqs:
            p = abi.Params()
            self.lib.ot2d_default_params(C.byref(p))
            p.update(**r.deck)
            params.append(p)
        mds = [abi.mesh_desc(r.grid.jmax, r.grid.kmax, grid.cgrid_bc(r.grid.jtail1)) for r in reqs]
        return abi.footprint(self.lib, "ot2d_", params, mds)

    def _slot_grid(self, r: Request):
        dims = tuple(self.slot_dims(r))
        if dims not in self._slot_grids:
            self._slot_grids[dims] = grid.naca4_cgrid(dims[0], dims[1], ds_wall=1e-6, wake_aspect=1000.0,
                                                      te_cluster=False)
        return self._slot_grids[dims]

    def cohort_footprint(self, r: Request) -> int:
        """Device bytes of a cohort of ``max_cohort`` slots of r's structure's
        largest grid."""
        g = self._slot_grid(r)
        n = self.max_cohort
        return abi.footprint(self.lib, "ot2d_", [_params(self.lib, r.deck)] * n, [_mesh(g)] * n)

    def run(self, requests: Sequence[Request], time_limit: Optional[float] = None,
            clock=time.perf_counter, progress=None) -> List[Result]:
        """Serve `requests` (each admitted no earlier than its arrival time)
        until all have finished or `time_limit` seconds have passed.
        `progress(dict)`, if given, is called after every tick with the
        counts of waiting, resident and finished requests."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        waiting: deque = deque()
        cohorts: List[_Cohort] = []
        results: Dict[int, Result] = {r.rid: Result(r.rid, r.kind, r.arrival, cells=r.cells) for r in requests}
        held = 0
        t0 = clock()
        while pending or waiting or cohorts:
            now = clock() - t0
            if time_limit is not None and now > time_limit:
                break
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # ---- cohorts whose enqueued steps are done: read their cases'
            # states (no wait: ot2d_query said the stream is idle)
            collected = False
            for c in cohorts:
                if c.busy and c.case.idle():
                    c.busy = False
                    self._collect(c, results, clock() - t0)
                    collected = True
            # ---- admission into the free slots of idle resident cohorts,
            # arrival order, each request into the smallest slot it fits
            still: deque = deque()
            resident = sum(len(c.open) for c in cohorts)
            while waiting:
                r = waiting.popleft()
                if self.max_resident is not None and resident >= self.max_resident:
                    still.append(r)
                    continue
                best = None
                for c in cohorts:
                    if c.busy:
                        continue
                    for i in c.free:
                        if c.fits(i, r):
                            area = c.cap[i][0] * c.cap[i][1]
                            if best is None or area < best[0]:
                                best = (area, c, i)
                if best is None:
                    still.append(r)
                    continue
                _, c, i = best
                c.admit(i, r)
                resident += 1
                results[r.rid].admitted = clock() - t0
                self.slot_admissions += 1
            waiting = still
            # cohorts with no case left running give their memory back (with
            # pre-sized slots, once no request is still to arrive)
            for c in [c for c in cohorts if not c.busy and not c.open and (self.slot_dims is None or not pending)]:
                held -= c.bytes
                c.close()
                cohorts.remove(c)
            # ---- new cohorts: arrival order, while the footprint fits.  A
            # request that fits a free slot of a cohort still running its
            # steps waits for that slot instead (one request per slot)
            spare = [(c, i) for c in cohorts if c.busy for i in c.free]
            admitted: Dict[tuple, List[Request]] = {}
            resident = sum(len(c.open) for c in cohorts)
            remaining: deque = deque()
            blocked = False
            for r in waiting:
                if blocked:
                    remaining.append(r)
                    continue
                slot = next((k for k, (c, i) in enumerate(spare) if c.fits(i, r)), None)
                if slot is not None:
                    spare.pop(slot)
                    remaining.append(r)
                    continue
                group = admitted.get(r.key(), [])
                if self.slot_dims is None:
                    need = self.footprint(group + [r]) - (self.footprint(group) if group else 0)
                else:
                    need = 0 if group else self.cohort_footprint(r)
                if ((self.max_resident is not None and resident >= self.max_resident)
                        or len(group) >= self.max_cohort or held + need > self.budget):
                    blocked = True
                    remaining.append(r)
                    continue
                held += need
                resident += 1
                admitted.setdefault(r.key(), []).append(r)
            waiting = remaining
            for group in admitted.values():
                if self.slot_dims is None:
                    c = _Cohort(self.lib, group, self.tol, self.errfn, self.tune_steps)
                    held += c.bytes - self.footprint(group)   # the case's own count
                else:
                    c = _Cohort(self.lib, group, self.tol, self.errfn, self.tune_steps, slots=self.max_cohort,
                                slot_grid=self._slot_grid(group[0]))
                    held += c.bytes - self.cohort_footprint(group[0])
                self.cohorts_made += 1
                tnow = clock() - t0
                for r in group:
                    results[r.rid].admitted = tnow
                cohorts.append(c)
            self.peak_bytes = max(self.peak_bytes, held)
            # ---- every idle cohort with a running case takes its next
            # steps_per_tick steps at once: cohorts advance independently,
            # their kernels overlapping on the device (each has its own stream)
            launched = False
            for c in cohorts:
                if not c.busy and c.open:
                    if c.dirty:
                        # launch over the running cases only (finished and
                        # empty slots cost no workgroups)
                        c.case.batch_compact()
                        c.dirty = False
                    c.case.iterate_async(self.k)
                    c.busy = True
                    launched = True
                    self.ticks += 1
            if (launched or collected) and progress is not None:
                progress(dict(t=clock() - t0, ticks=self.ticks, pending=len(pending), waiting=len(waiting),
                              resident=sum(len(c.open) for c in cohorts), cohorts=len(cohorts),
                              finished=sum(1 for x in results.values() if x.finished == x.finished)))
            if not pending and not waiting:
                # nothing left to take the free slots
                for c in [c for c in cohorts if not c.busy and not c.open]:
                    held -= c.bytes
                    c.close()
                    cohorts.remove(c)
            if not launched and not collected:
                if any(c.busy for c in cohorts):
                    time.sleep(0.001)
                elif pending:
                    time.sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
        for c in cohorts:
            c.close()
        return [results[r.rid] for r in requests]

    def _collect(self, c: "_Cohort", results: Dict[int, Result], tnow: float):
        """Per-case states of a cohort whose steps are done: a case that
        stopped (converged, failed) or reached max_steps finishes, and its
        slot becomes free."""
        st, ist = c.case.batch_status()
        for i in list(c.open):
            res = results[c.reqs[i].rid]
            res.steps = int(ist[i])
            if st[i] != abi.OK or ist[i] >= self.max_steps:
                if st[i] == abi.OK:
                    c.case.batch_retire(i)
                    res.status = -100          # not converged within max_steps
                else:
                    res.status = int(st[i])
                res.finished = tnow
                n = c.reqs[i].nspec
                res.forces = np.array([c.case.forces(i, s)[k] for s in range(n)
                                       for k in ("cl", "cd", "cm", "cpower", "cfx", "cfy")]).reshape(n, 6)
                if self.keep_state:
                    res.state = c.case.get_state(i)
                c.open.discard(i)
                c.free.add(i)
                c.dirty = True


# ---------------------------------------------------------------- streams
def polar_request(rid, arrival, rng, scale=1.0, lu_fp32=True) -> Request:
    """A steady polar point of the 64-case mix's distributions: NACA m~U[0,.06],
    p~U[.2,.6], t~U[.08,.18]; jmax~U{257..769}, kmax~U{65..193} (scaled);
    M~U[.3,.8], alpha~U[-4,8], Re log-U[1e6,2e7]; SA."""
    m, p, t = rng.uniform(0, 0.06), rng.uniform(0.2, 0.6), rng.uniform(0.08, 0.18)
    jmax = int(rng.integers(257, 770) * scale) // 2 * 2 + 1
    kmax = int(rng.integers(65, 194) * scale)
    fsmach, alfa = rng.uniform(0.3, 0.8), rng.uniform(-4, 8)
    rey = float(np.exp(rng.uniform(np.log(1e6), np.log(2e7))))
    g = grid.naca4_cgrid(jmax, kmax, m=m, p=p, t=t, ds_wall=1e-6, wake_aspect=1000.0, te_cluster=False)
    deck = dict(fsmach=fsmach, alfa=alfa, rey=rey, invisc=0, lamin=0, iturb=1, vnuinf=3.0, irhsy=-3, ilim=1,
                ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1, itnmax=1, timeac=1.0,
                lu_fp32=int(lu_fp32))
    return Request(rid, arrival, deck, g, "polar")


def pitch_request(rid, arrival, rng, scale=1.0, lu_fp32=True) -> Request:
    """A time-spectral pitching case: alpha(t) = alpha0 + alpha1 sin(omega t)
    with alpha0~U[0,5], alpha1~U[1,5] deg, reduced frequency k~U[.02,.1],
    N~{3,5,7} instances, 385x97 C-grid (scaled), M 0.3, Re 4e6, SA.  The
    reference's motion is alfa + angmax (1 - cos(rotf t)) (src/setup.f:367),
    so alfa = alpha0 - alpha1, angmax = alpha1 (a phase shift), rf = 2 k M."""
    a0, a1, k = rng.uniform(0, 5), rng.uniform(1, 5), rng.uniform(0.02, 0.1)
    n = int(rng.choice([3, 5, 7]))
    jmax = int(385 * scale) // 2 * 2 + 1
    kmax = int(97 * scale)
    g = grid.naca4_cgrid(jmax, kmax, ds_wall=1e-6, wake_aspect=1000.0, te_cluster=False)
    deck = dict(fsmach=0.3, alfa=a0 - a1, angmax=a1, rf=2 * k * 0.3, rey=4e6, invisc=0, lamin=0, iturb=1,
                vnuinf=3.0, irhsy=-3, ilim=1, ilhs=1, iconstant_cfl=1, dt=10.0, epse=0.01, iunst=0, ntac=1,
                itnmax=1, timeac=1.0, timespectral=1, nspec=n, motiontype=1, lu_fp32=int(lu_fp32))
    return Request(rid, arrival, deck, g, "pitch")


def stream_slot_dims(scale=1.0):
    """The largest grid of each structure in ``poisson_stream``: polar points
    (jmax up to 769, kmax up to 193) and pitching cases (385x97), scaled."""
    polar = (int(769 * scale) // 2 * 2 + 1, int(193 * scale))
    pitch = (int(385 * scale) // 2 * 2 + 1, int(97 * scale))
    return lambda r: pitch if r.kind == "pitch" else polar


def poisson_stream(lam, duration, seed=20261202, pitch_frac=0.3, scale=1.0, lu_fp32=True) -> List[Request]:
    """Poisson arrivals at ``lam`` requests/s for ``duration`` s."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / lam)
        if t > duration:
            break
        mk = pitch_request if rng.random() < pitch_frac else polar_request
        out.append(mk(len(out), t, rng, scale, lu_fp32))
    return out


def summarize(results: Sequence[Result], wall: float) -> dict:
    """Latency (arrival to convergence), its queueing (arrival to admission)
    and service (admission to convergence) parts over the converged
    requests, p50/p99; throughput over the wall time; and how many requests
    were still waiting or running when the loop ended."""
    ok = [r for r in results if r.status == abi.CONVERGED]

    def pct(v, q):
        return float(np.percentile(np.array(v), q)) if v else math.nan

    lat = [r.latency for r in ok]
    wait = [r.admitted - r.arrival for r in ok]
    svc = [r.finished - r.admitted for r in ok]
    cell_iters = float(sum(r.cells * r.steps for r in results))
    return dict(requests=len(results), converged=len(ok), wall_s=wall,
                p50_s=pct(lat, 50), p99_s=pct(lat, 99),
                queue_p50_s=pct(wait, 50), queue_p99_s=pct(wait, 99),
                service_p50_s=pct(svc, 50), service_p99_s=pct(svc, 99),
                steps_p50=pct([r.steps for r in ok], 50),
                converged_by_kind={k: sum(1 for r in ok if r.kind == k) for k in ("polar", "pitch")},
                never_admitted=sum(1 for r in results if r.admitted != r.admitted),
                failures=[dict(rid=r.rid, kind=r.kind, status=r.status, steps=r.steps) for r in results
                          if r.finished == r.finished and r.status not in (abi.CONVERGED, -100)],
                unfinished=sum(1 for r in results if r.admitted == r.admitted and r.finished != r.finished),
                solves_per_hr=len(ok) / wall * 3600.0 if wall > 0 else 0.0,
                cell_iter_per_s=cell_iters / wall if wall > 0 else 0.0)


def main(argv=None):
    import argparse
    import turns2d_urans_git_amd as t2d
    ap = argparse.ArgumentParser(description="serve a Poisson stream of polar and pitching requests")
    ap.add_argument("--rate", type=float, default=1.0)
    ap.add_argument("--duration", type=float, default=10.0)
    ap.add_argument("--scale", type=float, default=1.0, help="grid size factor")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--steps-per-tick", type=int, default=100)
    ap.add_argument("--max-steps", type=int, default=40000)
    ap.add_argument("--time-limit", type=float, default=None)
    ap.add_argument("--tune-steps", type=int, default=5, help="autotune steps per candidate (0: defaults)")
    a = ap.parse_args(argv)
    reqs = poisson_stream(a.rate, a.duration, scale=a.scale)
    srv = Server(t2d.load_library(), steps_per_tick=a.steps_per_tick, tol=a.tol, max_steps=a.max_steps,
                 errfn=t2d.last_error, tune_steps=a.tune_steps)
    t0 = time.perf_counter()
    res = srv.run(reqs, time_limit=a.time_limit)
    out = summarize(res, time.perf_counter() - t0)
    out.update(rate=a.rate, duration=a.duration, scale=a.scale, tol=a.tol, peak_bytes=srv.peak_bytes,
               cohorts_made=srv.cohorts_made, slot_admissions=srv.slot_admissions)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
