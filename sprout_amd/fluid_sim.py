"""FluidSim — the reference's solver-host interface (src/sim/mod.rs:862-3492) over the C ABI
of libsprout_hip.so.

Method names, argument meaning and error behaviour follow the reference's `FluidSim`:
`new`, `reset`, `set_solids`, `spawn`, `record`, `step_blocking`, `advance_blocking`,
`sim_time`, `download_particles`, `download_grid`, `download_pressure`, `cell_counts`,
`request_stats` / `stats`, `solver` (the SolverStats), `params`. Host buffers are copied in
and out; device memory stays owned by the handle.
"""
import ctypes as C

import numpy as np

from . import _native as N


class FluidSim:
    def __init__(self, grid, capacity, device=0, params=None):
        """FluidSim::new(gpu, grid, capacity) (mod.rs:1034)."""
        self.grid = (int(grid[0]), int(grid[1]))
        self.capacity = int(capacity)
        L = N.lib()
        cfg = N.Config(self.grid[0], self.grid[1], self.capacity, device)
        h = C.c_void_p()
        rc = L.sp_create(C.byref(cfg), C.byref(h))
        if rc != N.OK:
            raise N.SproutError(rc, L.sp_create_error().decode())
        self._h = h
        self._lib = L
        self.params = params if params is not None else N.default_params()
        self.set_params(self.params)

    # -- lifetime ----------------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.sp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, what):
        if rc != N.OK:
            raise N.SproutError(rc, f"{what}: {self._lib.sp_last_error(self._h).decode()}")

    # -- residency -------------------------------------------------------------------
    def evict(self):
        """Moves the tank's device state to host memory; the next call that needs it brings it back."""
        self._check(self._lib.sp_evict(self._h), "evict")

    @property
    def resident(self):
        r = C.c_int32()
        self._check(self._lib.sp_is_resident(self._h, C.byref(r)), "is_resident")
        return bool(r.value)

    # -- inputs ------------------------------------------------------------------------
    def set_params(self, params):
        """Assigning FluidSim::params (the grid-unit tables follow at the next record)."""
        self.params = params
        self._check(self._lib.sp_set_params(self._h, C.byref(params)), "set_params")

    def set_solids(self, field):
        """FluidSim::set_solids(SolidField) (mod.rs:1977)."""
        W, H = self.grid
        sdf = np.ascontiguousarray(field.sdf, np.float32)
        solid = np.ascontiguousarray(field.solid, np.float32)
        region = np.ascontiguousarray(field.region, np.uint32)
        assert sdf.size == (W + 1) * (H + 1) and solid.size == 4 * W * H and region.size == W * H
        temp = getattr(field, "temp", None)
        if temp is not None and len(temp):
            temp = np.ascontiguousarray(temp, np.float32)
            assert temp.size == W * H
        else:
            temp = None
        self._check(self._lib.sp_set_solids(self._h, sdf.ctypes.data, solid.ctypes.data, region.ctypes.data,
                                            None if temp is None else temp.ctypes.data), "set_solids")

    def reset(self, particles):
        """FluidSim::reset(Vec<Particle>) (mod.rs:1982-2001)."""
        p = np.ascontiguousarray(particles, N.PARTICLE_DTYPE)
        self._check(self._lib.sp_reset(self._h, p.ctypes.data if len(p) else None, len(p)), "reset")

    def spawn(self, particles):
        """FluidSim::spawn (mod.rs:2027-2048): returns how many were accepted."""
        p = np.ascontiguousarray(particles, N.PARTICLE_DTYPE)
        n = C.c_uint32()
        self._check(self._lib.sp_spawn(self._h, p.ctypes.data if len(p) else None, len(p), C.byref(n)), "spawn")
        return n.value

    # -- stepping ----------------------------------------------------------------------
    def record(self, substeps, target, tool=None):
        """FluidSim::record (mod.rs:2400-2563): enqueue only."""
        t = C.byref(tool) if tool is not None else None
        self._check(self._lib.sp_record(self._h, substeps, float(target), t), "record")

    def wait(self):
        self._check(self._lib.sp_wait(self._h), "wait")

    def step_blocking(self, substeps, tool=None):
        """FluidSim::step_blocking (mod.rs:3316-3338): ticks of 1/240 s under `tool`."""
        t = C.byref(tool) if tool is not None else None
        self._check(self._lib.sp_step_blocking(self._h, substeps, t), "step_blocking")

    def advance_blocking(self, seconds, tool=None):
        """FluidSim::advance_blocking (mod.rs:3284-3312): `tool` acts in the first record."""
        t = C.byref(tool) if tool is not None else None
        self._check(self._lib.sp_advance_blocking(self._h, float(seconds), t), "advance_blocking")

    def sim_time(self):
        t = C.c_double()
        self._check(self._lib.sp_sim_time(self._h, C.byref(t)), "sim_time")
        return t.value

    @property
    def solver(self):
        """FluidSim::solver (SolverStats, mod.rs:334-353)."""
        s = N.SolverStats()
        self._check(self._lib.sp_poll(self._h, C.byref(s)), "poll")
        return s

    def poll_readbacks(self):
        return self.solver

    def timings(self):
        """sp_timings: ms per substep of [total, particles, grid, pressure, projection,
        viscosity solve, second pressure solve + projection], then the iterations of the
        latest pressure, viscosity and second pressure solves."""
        out = np.zeros(10, np.float32)
        self._check(self._lib.sp_timings(self._h, out.ctypes.data), "timings")
        return out

    def kernel_time(self, kernel=0):
        """(average ms per launch, launches) of a hot kernel in the last batch."""
        ms, n = C.c_float(), C.c_uint32()
        self._check(self._lib.sp_kernel_time(self._h, kernel, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value

    def sort(self):
        self._check(self._lib.sp_sort(self._h), "sort")

    def download_ledger(self):
        """FluidSim::download_ledger (mod.rs:3389-3406): energy that crossed the boundary since
        the last reset, J per metre of depth: gained from / lost to air, gained from / lost to
        solids, heat tool, cool tool, boiled off, made inside."""
        out = (C.c_double * 8)()
        self._check(self._lib.sp_download_ledger(self._h, out), "download_ledger")
        return [float(x) for x in out]

    def download_ledger_counts(self):
        """Per table row: (boiled or evaporated away, residues left, steam made) since the last
        reset (FluidSim::download_ledger's counts, mod.rs:3400-3404)."""
        out = [np.zeros(N.MAX_MATERIALS, np.uint32) for _ in range(3)]
        self._check(self._lib.sp_download_ledger_counts(self._h, *(o.ctypes.data for o in out)), "download_ledger_counts")
        return tuple(out)

    def set_magnets(self, magnets):
        """FluidSim::set_magnets (mod.rs:2018-2022): a list of N.Magnet."""
        arr = (N.Magnet * max(len(magnets), 1))(*magnets)
        self._check(self._lib.sp_set_magnets(self._h, arr, len(magnets)), "set_magnets")

    # -- render-facing outputs (mod.rs:1608-1943) ---------------------------------------
    def set_currents(self, currents):
        """FluidSim::set_currents (mod.rs:2008-2012): Current records (the first 16 kept)."""
        arr = (N.Current * max(len(currents), 1))(*currents)
        self._check(self._lib.sp_set_currents(self._h, arr, len(currents)), "set_currents")

    def render_view(self, which):
        """The device view (address, bytes, count, stride, device, generation) of a render
        buffer: "PARTICLES", "SDF", "CELL", "COMPOSITION", "SURFACE", "MATERIALS", "VELOCITY", "GPUMAT"."""
        v = N.DeviceView()
        self._check(self._lib.sp_render_view(self._h, N.RENDER[which], C.byref(v)), f"render_view {which}")
        return v

    def render_download(self, which):
        """A host copy of a render buffer, as bytes-shaped numpy array of its elements."""
        v = self.render_view(which)
        out = np.zeros(v.bytes, np.uint8)
        self._check(self._lib.sp_render_download(self._h, N.RENDER[which], out.ctypes.data if v.bytes else None, out.nbytes),
                    f"render_download {which}")
        return out

    def stream(self):
        h = C.c_void_p()
        self._check(self._lib.sp_stream(self._h, C.byref(h)), "stream")
        return h.value

    def resets(self):
        n = C.c_uint64()
        self._check(self._lib.sp_resets(self._h, C.byref(n)), "resets")
        return n.value

    def shares_a_group(self):
        out = C.c_int32()
        self._check(self._lib.sp_shares_a_group(self._h, C.byref(out)), "shares_a_group")
        return bool(out.value)

    # -- outputs -----------------------------------------------------------------------
    def download_particles(self):
        """FluidSim::download_particles (mod.rs:3421-3432): live particles in slot order."""
        n = C.c_uint32()
        self._check(self._lib.sp_download_particles(self._h, None, 0, C.byref(n)), "download_particles")
        out = np.zeros(n.value, N.PARTICLE_DTYPE)
        if n.value:
            self._check(self._lib.sp_download_particles(self._h, out.ctypes.data, n.value, C.byref(n)),
                        "download_particles")
        return out[: n.value]

    def download_grid(self):
        """FluidSim::download_grid (mod.rs:3436-3450): (vel[H,W,4], cell[H,W,4])."""
        W, H = self.grid
        vel = np.zeros((H, W, 4), np.float32)
        cell = np.zeros((H, W, 4), np.float32)
        self._check(self._lib.sp_download_grid(self._h, vel.ctypes.data, cell.ctypes.data), "download_grid")
        return vel, cell

    def download_labels(self):
        """The rigid `labels` of GridDump (mod.rs:3436-3450): [H,W,2] int32, per cell the root
        cell of its body (-1: not rigid) and, at a root, the body's slot."""
        W, H = self.grid
        lab = np.zeros((H, W, 2), np.int32)
        self._check(self._lib.sp_download_labels(self._h, lab.ctypes.data), "download_labels")
        return lab

    def download_bodies(self):
        """FluidSim::download_bodies (mod.rs:3451-3462): [bodies, 7] float64 (mass, offset sums
        x, y, momentum x, y, angular momentum about the root, inertia sum)."""
        n = C.c_uint32()
        self._check(self._lib.sp_download_bodies(self._h, None, 0, C.byref(n)), "download_bodies")
        out = np.zeros((n.value, 7), np.float64)
        if n.value:
            self._check(self._lib.sp_download_bodies(self._h, out.ctypes.data, n.value, C.byref(n)), "download_bodies")
        return out[:n.value]

    def slots(self):
        """FluidSim::slots (mod.rs:1946)."""
        v = C.c_uint32()
        self._check(self._lib.sp_slots(self._h, C.byref(v)), "slots")
        return v.value

    def dt_estimate(self):
        """FluidSim::dt_estimate (mod.rs:1961)."""
        v = C.c_float()
        self._check(self._lib.sp_dt_estimate(self._h, C.byref(v)), "dt_estimate")
        return v.value

    def recorded_iters(self):
        """FluidSim::recorded_iters (mod.rs:1966): [pressure, second pressure, viscosity]."""
        v = (C.c_uint32 * 3)()
        self._check(self._lib.sp_recorded_iters(self._h, v), "recorded_iters")
        return [int(x) for x in v]

    def substep_cost(self):
        """FluidSim::substep_cost (mod.rs:1972): device s per substep, smoothed; 0 before the
        first measured batch."""
        v = C.c_float()
        self._check(self._lib.sp_substep_cost(self._h, C.byref(v)), "substep_cost")
        return v.value

    def frame_blocking(self, substeps, target, tool=None):
        """FluidSim::frame_blocking (mod.rs:3342-3350): the simulated time covered, s."""
        v = C.c_double()
        t = C.byref(tool) if tool is not None else None
        self._check(self._lib.sp_frame_blocking(self._h, substeps, target, t, C.byref(v)), "frame_blocking")
        return v.value

    def profile_passes(self):
        """FluidSim::profile_passes (mod.rs:3120-3150): one substep, [(phase, ms)], with the
        total last."""
        out = (N.PassTime * 16)()
        n = C.c_uint32()
        self._check(self._lib.sp_profile_passes(self._h, out, 16, C.byref(n)), "profile_passes")
        return [(out[k].name.decode(), out[k].ms) for k in range(n.value)]

    def material_table(self):
        """The GpuMat rows of material_buffer (material.rs:508-549) as a structured array."""
        out = np.zeros(N.MAX_MATERIALS, N.GPUMAT_DTYPE)
        self._check(self._lib.sp_material_table(self._h, out.ctypes.data, N.MAX_MATERIALS), "material_table")
        return out

    def download_pressure(self):
        """FluidSim::download_pressure (mod.rs:3464-3475): (p*dt per cell [H,W], dt)."""
        W, H = self.grid
        p = np.zeros((H, W), np.float32)
        dt = C.c_float()
        self._check(self._lib.sp_download_pressure(self._h, p.ctypes.data, C.byref(dt)), "download_pressure")
        return p, dt.value

    def cell_counts(self, lo, hi):
        """FluidSim::cell_counts(lo, hi) (mod.rs:3354-3386)."""
        w, h = hi[0] - lo[0], hi[1] - lo[1]
        out = np.zeros(max(w, 0) * max(h, 0), np.uint32)
        self._check(self._lib.sp_cell_counts(self._h, lo[0], lo[1], w, h, out.ctypes.data if out.size else None),
                    "cell_counts")
        return out

    def stats(self):
        """Stats (mod.rs:128-143, stats.comp): live particles, per-row counts and concentrations."""
        s = N.StatsOut()
        self._check(self._lib.sp_stats(self._h, C.byref(s)), "stats")
        return s

    # -- stage-level access (parity tests, diagnostics) ---------------------------------
    def field(self, name, dtype, shape):
        out = np.zeros(shape, dtype)
        self._check(self._lib.sp_debug_download(self._h, N.FIELD[name], out.ctypes.data, out.nbytes), f"field {name}")
        return out

    def put_field(self, name, arr):
        arr = np.ascontiguousarray(arr)
        self._check(self._lib.sp_debug_upload(self._h, N.FIELD[name], arr.ctypes.data, arr.nbytes), f"put {name}")

    def stage(self, name, arg=0):
        self._check(self._lib.sp_debug_stage(self._h, N.STAGE[name], arg), f"stage {name}")

    def grid_params(self):
        g = N.GParams()
        mats = (N.GMat * N.MAX_MATERIALS)()
        self._check(self._lib.sp_debug_gparams(self._h, C.byref(g), mats), "grid_params")
        return g, mats


class FluidPool:
    """Continuous batching of step requests over many resident scenes (sp_pool_*), run in
    rounds by the thread that calls step(), wait() or drain(); submit() may be called from
    any thread. The FluidSim handles added belong to the pool until close(); each scene's
    results are those it gets stepped alone with the same requests."""

    def __init__(self, budget_bytes, streams=4, device=0):
        L = N.lib()
        h = C.c_void_p()
        rc = L.sp_pool_create(device, int(budget_bytes), streams, C.byref(h))
        if rc != N.OK:
            raise N.SproutError(rc, "sp_pool_create")
        self._h = h
        self._lib = L
        self.sims = []

    def close(self):
        """Runs what was submitted and gives the handles back."""
        if getattr(self, "_h", None):
            self._lib.sp_pool_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != N.OK:
            raise N.SproutError(rc, f"{what}: {self._lib.sp_pool_last_error(self._h).decode()}")

    def add(self, sim):
        """Adds a set-up FluidSim as the next scene; returns its id."""
        k = C.c_uint32()
        self._check(self._lib.sp_pool_add(self._h, sim._h, C.byref(k)), "pool add")
        self.sims.append(sim)
        return k.value

    def submit(self, scene, substeps, target):
        t = C.c_uint64()
        self._check(self._lib.sp_pool_submit(self._h, scene, substeps, float(target), C.byref(t)), "pool submit")
        return t.value

    def step(self, wait_ms=0.0):
        """One round; returns the scenes it advanced."""
        n = C.c_uint32()
        self._check(self._lib.sp_pool_step(self._h, float(wait_ms), C.byref(n)), "pool step")
        return n.value

    def wait(self, ticket):
        info = N.RequestInfo()
        self._check(self._lib.sp_pool_wait(self._h, ticket, C.byref(info)), "pool wait")
        return info

    def poll(self, ticket):
        """The request's info, or None while it is not done."""
        info = N.RequestInfo()
        rc = self._lib.sp_pool_poll(self._h, ticket, C.byref(info))
        if rc == N.E_BUSY:
            return None
        self._check(rc, "pool poll")
        return info

    def drain(self):
        self._check(self._lib.sp_pool_drain(self._h), "pool drain")

    def stats(self):
        st = N.PoolStats()
        self._check(self._lib.sp_pool_stats_get(self._h, C.byref(st)), "pool stats")
        return {name: getattr(st, name) for name, _ in N.PoolStats._fields_}

    def reset_stats(self):
        self._check(self._lib.sp_pool_reset_stats(self._h), "pool reset_stats")


class FluidBatch:
    """Many FluidSim handles advanced together (sp_batch_*): tanks of the same grid and
    solver schedule share each kernel launch of a substep, groups overlap on the device, and
    each tank's results are those it gets stepped alone."""

    def __init__(self, sims):
        self.sims = list(sims)
        L = N.lib()
        arr = (C.c_void_p * len(self.sims))(*[s._h.value for s in self.sims])
        h = C.c_void_p()
        rc = L.sp_batch_create(arr, len(self.sims), C.byref(h))
        if rc != N.OK:
            raise N.SproutError(rc, "sp_batch_create")
        self._h = h
        self._lib = L

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sp_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc, what):
        if rc != N.OK:
            raise N.SproutError(rc, f"{what}: {self._lib.sp_batch_last_error(self._h).decode()}")

    def record(self, substeps, target):
        self._check(self._lib.sp_batch_record(self._h, substeps, float(target)), "batch record")

    def wait(self):
        self._check(self._lib.sp_batch_wait(self._h), "batch wait")

    def groups(self):
        n = C.c_uint32()
        sizes = np.zeros(len(self.sims), np.uint32)
        self._check(self._lib.sp_batch_groups(self._h, C.byref(n), sizes.ctypes.data, len(sizes)), "batch groups")
        return sizes[: n.value].tolist()

    def timings(self, group):
        out = np.zeros(8, np.float32)
        self._check(self._lib.sp_batch_timings(self._h, group, out.ctypes.data), "batch timings")
        return out
