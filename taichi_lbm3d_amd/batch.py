"""Batched solver: many porous samples of uneven size stepped together.

The reference runs one grid per solver object (SP:9-114) and its drivers loop
over samples in Python (example_porous_medium.py).  Here the fluid lists of
all samples are laid end to end in one population array and every time step
of the whole batch is one kernel launch (lbm3d_batch_step): a persistent grid
of workgroups walks tiles of cells, each tile inside one sample, whose grid
it reads from a per-sample descriptor.  Each sample evolves exactly as
LB3D_Solver_Single_Phase would evolve it alone (bit-identical: same per-cell
code, same arithmetic).

Samples share the physical parameters (niu, force, convention, storage type)
and the face modes; each has its own grid.  Converged samples can be retired:
their cells are skipped by later launches.  The step loop can be captured in a
HIP graph (torch.cuda.CUDAGraph) and replayed; launch shapes come from the
autotuner in autotune.py.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .solver import CONVENTIONS, FACE_NAMES, relaxation_rates, _ptr, _ref, _torch, ctypes_void


DEFAULT_GROUP_BYTES = 128 << 20  # population bytes per group of the cache-blocked schedule


def group_ranges(begins, total: int, bytes_per_cell: int, group_bytes: int):
    """[(cell_begin, cell_end)]: consecutive samples (first cells `begins`,
    increasing, the batch ending at `total`) gathered into runs whose cells
    take at most group_bytes; a sample larger than that forms a run of its own."""
    bounds = list(begins) + [int(total)]
    groups, start = [], bounds[0]
    for s in range(1, len(bounds)):
        if (bounds[s] - start) * bytes_per_cell > group_bytes and bounds[s - 1] > start:
            groups.append((start, bounds[s - 1]))
            start = bounds[s - 1]
    groups.append((start, bounds[-1]))
    return groups


class BatchSolver:
    """Batch of samples with shared physics.

    solids      -- sequence of int8/uint8 arrays (nx, ny, nz); a cell is fluid iff 0
    niu, force  -- as the reference's niu and (fx, fy, fz) (SP:13-17)
    faces       -- optional {face name: ("pressure", rho) | ("velocity", (vx, vy, vz))},
                   face names as FACE_NAMES; unnamed faces are periodic
    convention, dtype, device, band_rows -- as LB3D_Solver_Single_Phase
    """

    def __init__(self, solids: Sequence[np.ndarray], *, niu: float = 0.16667, force=(0.0, 0.0, 0.0),
                 faces: Optional[dict] = None, convention: str = "reference", dtype: str = "f32", device=None,
                 use_tiles: bool = False, group_bytes: Optional[int] = None, steps_per_visit: int = 32,
                 band_rows: int = 32):
        if not solids:
            raise ValueError("a batch needs at least one sample")
        if convention not in CONVENTIONS:
            raise ValueError(f"unknown convention {convention!r}")
        if dtype not in _lib.DTYPES:
            raise ValueError(f"dtype must be one of {sorted(_lib.DTYPES)}")
        self.solids = [np.ascontiguousarray(np.asarray(s).astype(np.int8, copy=False)) for s in solids]
        for s in self.solids:
            if s.ndim != 3:
                raise ValueError("every sample must be a 3-D array")
        self.niu = float(niu)
        self.fx, self.fy, self.fz = (float(f) for f in force)
        self.convention = convention
        self.dtype = dtype
        self.faces = dict(faces or {})
        for name in self.faces:
            if name not in FACE_NAMES:
                raise ValueError(f"unknown face {name!r}; faces are {FACE_NAMES}")
        torch = _torch()
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
        self.n_samples = len(self.solids)
        self._t = 0
        self.launch = None  # (even, odd) lbm3d_launch shapes; None = library defaults
        # use_tiles: odd steps with 256-thread blocks locate their cells through
        # tile descriptors (LBM3D_FLAG_TILES); identical results, but slower in
        # a batch (the descriptor load waits for the sample lookup), so off
        self.use_tiles = bool(use_tiles)
        # order of every sample's fluid list, as LB3D_Solver_Single_Phase.band_rows
        # (the same default, so a sample's slots match its single-grid run's)
        self.band_rows = int(band_rows)
        # Cache-blocked schedule: with group_bytes > 0 and steps_per_visit > 0,
        # consecutive samples are grouped into runs of at most group_bytes of
        # population state, and run() advances one group by steps_per_visit
        # steps before it moves to the next, so each group's slots stay in the
        # 256 MB last-level cache across its steps.  Samples are independent,
        # so the results are the same bit for bit.  group_bytes=None: groups
        # of 128 MiB once the whole state exceeds twice that, else one group.
        self._group_bytes_arg = group_bytes
        self.steps_per_visit = int(steps_per_visit)
        self._graph = None
        self._build()
        per_cell = 19 * _lib.torch_storage_dtype(self.dtype).itemsize
        if group_bytes is None:
            group_bytes = DEFAULT_GROUP_BYTES if int(self.sizes.cells) * per_cell > 2 * DEFAULT_GROUP_BYTES else 0
        self.group_bytes = int(group_bytes)
        self.groups = self._make_groups()
        self._live = list(range(len(self.groups)))

    # ---- set-up ---------------------------------------------------------------
    def _params(self):
        p = _lib.Params()
        _, sv, so = relaxation_rates(self.niu, self.convention)
        p.s_v, p.s_o = sv, so
        p.guo_a, p.guo_b = CONVENTIONS[self.convention]
        p.fx, p.fy, p.fz = self.fx, self.fy, self.fz
        p.force_on = 1 if (self.fx != 0 or self.fy != 0 or self.fz != 0) else 0
        p.flags = _lib.STORAGE_FLAGS.get(self.dtype, 0)
        p.band_rows = self.band_rows
        for f, name in enumerate(FACE_NAMES):
            p.face_u_off[f] = -1
            p.rho_bc[f] = 1.0
            spec = self.faces.get(name)
            if spec is None:
                p.bc_mode[f] = 0
            elif spec[0] == "pressure":
                p.bc_mode[f] = 1
                p.rho_bc[f] = float(spec[1])
            elif spec[0] == "velocity":
                p.bc_mode[f] = 2
                for c in range(3):
                    p.u_bc[f][c] = float(spec[1][c])
            else:
                raise ValueError(f"face {name}: mode must be 'pressure' or 'velocity'")
        return p

    def _sample_params(self, s: int):
        q = _lib.Params()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(self.p), ctypes.sizeof(q))
        q.nx, q.ny, q.nz = self.solids[s].shape
        q.n_cells = self.samples[s].n_cells if hasattr(self, "samples") else 0
        return q

    def _stream(self):
        torch = _torch()
        return ctypes_void(torch.cuda.current_stream(self.device).cuda_stream)

    def _build(self):
        torch = _torch()
        lib = _lib.lib()
        dev = self.device
        self.p = self._params()
        n = self.n_samples
        self.samples = (_lib.Sample * n)()
        solids_d, counts = [], torch.zeros(n, dtype=torch.int32, device=dev)
        with torch.cuda.device(dev):
            st = self._stream()
            ws_bytes = max(lib.size("lbm3d_compact_workspace_bytes", int(np.prod(s.shape))) for s in self.solids)
            wss = []
            for s in range(n):
                q = self._sample_params(s)
                sd = torch.from_numpy(self.solids[s].reshape(-1)).to(dev)
                ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
                lib.call("lbm3d_compact_count", st, _ref(q), _ptr(sd), ctypes_void(counts.data_ptr() + 4 * s),
                         _ptr(ws), ws_bytes)
                solids_d.append(sd)
                wss.append(ws)
            cnt = counts.cpu().numpy()
            for s in range(n):
                self.samples[s].nx, self.samples[s].ny, self.samples[s].nz = self.solids[s].shape
                self.samples[s].n_cells = int(cnt[s])
            sizes = _lib.BatchSizes()
            lib.call("lbm3d_batch_layout", _ref(self.p), self.samples, n, ctypes.byref(sizes))
            self.sizes = sizes
            self.p.n_cells = int(sizes.cells)
            total = int(sizes.cells)
            cells = torch.full((total,), -1, dtype=torch.int32, device=dev)
            n_tiles = total // _lib.TILE_CELLS if self.use_tiles else 0
            rows = torch.empty(max(4 * (int(sizes.row_words) + n_tiles), 4), dtype=torch.int32, device=dev)
            face_bits = torch.zeros(lib.size("lbm3d_face_bit_words", total), dtype=torch.int32, device=dev)
            for s in range(n):
                sm = self.samples[s]
                q = self._sample_params(s)
                d2c = torch.empty(int(np.prod(self.solids[s].shape)), dtype=torch.int32, device=dev)
                cb = int(sm.cell_begin)
                lib.call("lbm3d_compact_fill", st, _ref(q), _ptr(solids_d[s]), _ptr(d2c),
                         ctypes_void(cells.data_ptr() + 4 * cb), _ptr(wss[s]), ws_bytes)
                lib.call("lbm3d_build_rows_at", st, _ref(q), _ptr(solids_d[s]), _ptr(d2c), cb,
                         ctypes_void(rows.data_ptr() + 16 * int(sm.rows_begin)))
                lib.call("lbm3d_build_face_bits", st, _ref(q), _ptr(solids_d[s]),
                         ctypes_void(cells.data_ptr() + 4 * cb), ctypes_void(face_bits.data_ptr() + 4 * (cb // 16)))
                if n_tiles:
                    lib.call("lbm3d_build_tiles", st, _ref(q), ctypes_void(cells.data_ptr() + 4 * cb),
                             ctypes_void(rows.data_ptr() + 16 * (int(sizes.row_words) + cb // _lib.TILE_CELLS)))
                del d2c
            if n_tiles:
                self.p.flags |= _lib.FLAG_TILES
            else:
                self.p.flags &= ~_lib.FLAG_TILES
            host_desc = np.zeros(int(sizes.desc_bytes), np.uint8)
            lib.call("lbm3d_batch_desc", _ref(self.p), self.samples, n, ctypes_void(host_desc.ctypes.data),
                     int(sizes.desc_bytes))
            desc = torch.from_numpy(host_desc).to(dev)
            tdt = _lib.torch_storage_dtype(self.dtype)
            stride = lib.size("lbm3d_slot_stride", total)
            g = torch.zeros(19 * stride, dtype=tdt, device=dev)
            face_u = torch.zeros(max(int(sizes.face_u), 1), dtype=self._real_dtype(), device=dev)
            active = torch.ones(n, dtype=torch.uint8, device=dev)
            torch.cuda.current_stream(dev).synchronize()
        self._dev = dict(cells=cells, rows=rows, face_bits=face_bits, desc=desc, g=g, face_u=face_u, active=active)
        self._init_state()

    def _real_dtype(self):
        torch = _torch()
        return torch.float64 if self.dtype == "f64" else torch.float32

    def _init_state(self, rho0=None, u0=None):
        torch = _torch()
        d = self._dev
        with torch.cuda.device(self.device):
            _lib.lib().call("lbm3d_batch_init", self._stream(), _ref(self.p), _lib.DTYPES[self.dtype], _ptr(d["g"]),
                            _ptr(d["desc"]), self.n_samples, _ptr(d["cells"]), _ptr(d["face_u"]),
                            _ptr(rho0) if rho0 is not None else None, _ptr(u0) if u0 is not None else None)
        self._t = 0
        self._graph = None
        d["active"].fill_(1)
        # every sample active: the steps get no mask (NULL), which lets even
        # steps load populations before looking their sample up
        self._all_active = True
        if hasattr(self, "groups"):
            self._live = list(range(len(self.groups)))

    def init_from_macros(self, fields):
        """Warm start: f = feq(rho, v) per sample from dense (rho (nx,ny,nz), v (nx,ny,nz,3))
        pairs (numpy arrays or torch tensors)."""
        torch = _torch()
        total = int(self.sizes.cells)
        rho = torch.ones(total, dtype=torch.float32, device=self.device)
        u = torch.zeros(3, total, dtype=torch.float32, device=self.device)
        cells = self._dev["cells"]
        for s, (r, v) in enumerate(fields):
            sm = self.samples[s]
            cb, n = int(sm.cell_begin), int(sm.n_cells)
            idx = cells[cb:cb + n].long()
            rd = (r if torch.is_tensor(r) else torch.from_numpy(np.asarray(r))).to(self.device, torch.float32)
            vd = (v if torch.is_tensor(v) else torch.from_numpy(np.asarray(v))).to(self.device, torch.float32)
            rd, vd = rd.reshape(-1), vd.reshape(-1, 3)
            rho[cb:cb + n] = rd[idx]
            u[:, cb:cb + n] = vd[idx].t()
        self._init_state(rho, u.contiguous())

    # ---- stepping -----------------------------------------------------------------
    @property
    def steps_done(self) -> int:
        return self._t

    @property
    def n_fluid(self):
        return np.array([int(self.samples[s].n_cells) for s in range(self.n_samples)])

    def _launch_arg(self):
        if self.launch is None:
            return None
        arr = (_lib.Launch * 2)()
        for i, c in enumerate(self.launch):
            arr[i].block, arr[i].grid_per_cu, arr[i].reverse = int(c[0]), int(c[1]), int(c[2])
        return arr

    def set_schedule(self, group_bytes: int, steps_per_visit: int = 32):
        """Switch the cache-blocked schedule: groups of at most group_bytes, or
        one group (a launch per step over the whole batch) with group_bytes 0."""
        self.group_bytes, self.steps_per_visit = int(group_bytes), int(steps_per_visit)
        self.groups = self._make_groups()
        keep = self._dev["active"].cpu().numpy().astype(bool)
        begins = np.array([int(self.samples[s].cell_begin) for s in range(self.n_samples)])
        self._live = [i for i, (cb, ce) in enumerate(self.groups) if keep[(begins >= cb) & (begins < ce)].any()]
        self._graph = None

    def _make_groups(self):
        """Cell ranges of the cache-blocked schedule (group_ranges); one group when off."""
        total = int(self.sizes.cells)
        if self.group_bytes <= 0 or self.steps_per_visit <= 0:
            return [(0, total)]
        per_cell = 19 * _lib.torch_storage_dtype(self.dtype).itemsize
        begins = [int(self.samples[s].cell_begin) for s in range(self.n_samples)]
        return group_ranges(begins, total, per_cell, self.group_bytes)

    def _enqueue(self, nsteps: int, t0: int):
        d = self._dev
        lib = _lib.lib()
        args = (self._stream(), _ref(self.p), _lib.DTYPES[self.dtype], _ptr(d["g"]), _ptr(d["desc"]), self.n_samples,
                _ptr(d["cells"]), _ptr(d["rows"]), _ptr(d["face_bits"]), _ptr(d["face_u"]),
                None if self._all_active else _ptr(d["active"]))
        if len(self.groups) == 1:
            lib.call("lbm3d_batch_run", *args, int(t0), int(nsteps), self._launch_arg())
            return
        live = [self.groups[i] for i in self._live]  # groups with an active sample
        done, sweep = 0, 0
        while done < nsteps:
            m = min(self.steps_per_visit, nsteps - done)
            # alternate the group order, so a sweep starts with the group the last one ended on
            order = live if sweep % 2 == 0 else live[::-1]
            for cb, ce in order:
                lib.call("lbm3d_batch_run_range", *args, int(t0 + done), int(m), int(cb), int(ce),
                         self._launch_arg())
            done += m
            sweep += 1

    def set_launch(self, even, odd):
        """Launch shapes (block, grid_per_cu, reverse) of even and odd steps."""
        self.launch = (tuple(even), tuple(odd))
        self._graph = None

    def step(self):
        self.run(1)

    def run(self, nsteps: int, graph_steps: int = 0):
        """Enqueue nsteps steps of every active sample.  With graph_steps (even, > 0)
        the steady part of the loop replays a captured HIP graph of that many steps."""
        torch = _torch()
        nsteps = int(nsteps)
        with torch.cuda.device(self.device):
            if graph_steps <= 0 or nsteps < graph_steps + 2:
                self._enqueue(nsteps, self._t)
                self._t += nsteps
                return
            if graph_steps % 2:
                raise ValueError("graph_steps must be even (the graph starts on an even step)")
            if self.p.flags & _lib.FLAG_STOCHASTIC_ROUNDING:
                # the step's rounding seed is a kernel argument: a replayed graph would
                # repeat the random bits of the steps it was captured at
                raise ValueError(f"graph_steps needs round-to-nearest storage; {self.dtype!r} rounds "
                                 "stochastically with per-step seeds that a replay would repeat")
            # eager steps until t is even and past the first step
            lead = 0
            while (self._t + lead) % 2 or self._t + lead == 0:
                lead += 1
            if lead:
                self._enqueue(lead, self._t)
                self._t += lead
                nsteps -= lead
            if self._graph is None or self._graph[0] != graph_steps:
                # two real eager steps (one per parity) load the kernels before the capture
                self._enqueue(2, self._t)
                self._t += 2
                nsteps -= 2
                self._capture(graph_steps)
            reps, rest = divmod(nsteps, graph_steps)
            g = self._graph[1]
            for _ in range(reps):
                g.replay()
            self._t += reps * graph_steps
            if rest:
                self._enqueue(rest, self._t)
                self._t += rest

    def _capture(self, graph_steps: int):
        """Capture graph_steps steps starting on an even step t >= 2: their
        kernel arguments are then the same for every replay at such a t."""
        torch = _torch()
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._enqueue(graph_steps, 2)
        self._graph = (graph_steps, g)

    # ---- fields and reductions ---------------------------------------------------------
    def macros(self):
        """FP32 (rho [cells], v [3*cells]) in batch ids after steps_done steps."""
        torch = _torch()
        d = self._dev
        total = int(self.sizes.cells)
        rho = torch.empty(total, dtype=self._real_dtype(), device=self.device)
        v = torch.empty(3 * total, dtype=self._real_dtype(), device=self.device)
        with torch.cuda.device(self.device):
            _lib.lib().call("lbm3d_batch_macros", self._stream(), _ref(self.p), _lib.DTYPES[self.dtype],
                            _ptr(d["g"]), _ptr(d["desc"]), self.n_samples, _ptr(d["cells"]), _ptr(d["rows"]),
                            _ptr(d["face_bits"]), _ptr(d["face_u"]), self._t, _ptr(rho), _ptr(v))
        if self.dtype == "f64":
            rho, v = rho.float(), v.float()
        return rho, v

    def reduce(self):
        """Per-sample FP64 sums: array (n_samples, 8) = sum rho, sum v (3), max |v|,
        non-finite cells, fluid cells, 0."""
        torch = _torch()
        rho, v = self.macros()
        lib = _lib.lib()
        ws_bytes = lib.size("lbm3d_batch_reduce_workspace_bytes", int(self.sizes.cells))
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        out = torch.zeros(self.n_samples * 8, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            lib.call("lbm3d_batch_reduce", self._stream(), _ref(self.p), _ptr(self._dev["desc"]), self.n_samples,
                     _ptr(rho), _ptr(v), _ptr(out), _ptr(ws), ws_bytes)
        return out.cpu().numpy().reshape(self.n_samples, 8)

    def mean_vx(self) -> np.ndarray:
        r = self.reduce()
        return r[:, 1] / np.maximum(r[:, 6], 1)

    def effective_viscosity(self) -> float:
        tau, _, _ = relaxation_rates(self.niu, self.convention)
        return (tau - 0.5) / 3.0

    def permeability(self, r=None) -> np.ndarray:
        """Per-sample Darcy permeability along x under the body force fx, lattice
        units^2: k = nu_eff q / F_eff with the momentum-consistent superficial
        velocity q (the formula of LB3D_Solver_Single_Phase.permeability);
        `r` an earlier reduce() result, or None to reduce now."""
        if self.fx == 0:
            raise ValueError("batch permeability needs a body force fx")
        nu = self.effective_viscosity()
        a, _ = CONVENTIONS[self.convention]
        f_eff = a * self.fx / 3.0
        if r is None:
            r = self.reduce()
        vol = np.array([float(np.prod(s.shape)) for s in self.solids])
        q = (r[:, 1] - 0.5 * (self.fx - f_eff) * r[:, 6]) / vol
        return nu * q / f_eff

    def sample_fields(self, s: int):
        """Dense (rho (nx,ny,nz), v (nx,ny,nz,3)) of sample s, as to_numpy() of the
        single solver with dense storage (non-fluid cells rho 1, v 0, SP:390-392)."""
        torch = _torch()
        rho, v = self.macros()
        sm = self.samples[s]
        q = self._sample_params(s)
        shape = self.solids[s].shape
        rho_d = torch.ones(shape, dtype=torch.float32, device=self.device)
        v_d = torch.zeros(shape + (3,), dtype=torch.float32, device=self.device)
        cb = int(sm.cell_begin)
        with torch.cuda.device(self.device):
            _lib.lib().call("lbm3d_scatter_dense", self._stream(), _ref(q),
                            ctypes_void(self._dev["cells"].data_ptr() + 4 * cb), ctypes_void(rho.data_ptr() + 4 * cb),
                            ctypes_void(v.data_ptr() + 4 * cb), int(self.sizes.cells), shape[0], shape[1], shape[2],
                            _ptr(rho_d), _ptr(v_d))
        return rho_d.cpu().numpy(), v_d.cpu().numpy()

    def retire(self, mask):
        """Stop stepping the samples where mask is true (their state stays as is)."""
        torch = _torch()
        keep = ~np.asarray(mask, bool)
        self._dev["active"].copy_(torch.from_numpy(keep.astype(np.uint8)).to(self.device))
        if self._all_active != bool(keep.all()):
            self._all_active = bool(keep.all())
            self._graph = None  # a captured graph holds the mask argument
        # groups whose samples have all retired are no longer launched
        begins = np.array([int(self.samples[s].cell_begin) for s in range(self.n_samples)])
        live = [i for i, (cb, ce) in enumerate(self.groups) if keep[(begins >= cb) & (begins < ce)].any()]
        if live != self._live:
            self._live = live
            self._graph = None

    def run_to_steady(self, tol: float = 1e-6, interval: int = 500, max_steps: int = 50_000, graph_steps: int = 0,
                      callback=None, check_every: int = 0):
        """Step until every sample meets |d<vx>| / |<vx>| < tol between checks
        `interval` steps apart; a sample that meets it is retired at once.
        Reductions run every `check_every` steps (default `interval`, which it
        must divide); callback(t, reduce_rows) is called after each.
        Returns (steps per sample, converged mask)."""
        check_every = check_every or interval
        if interval % 2 or check_every % 2 or interval % check_every:
            raise ValueError("interval and check_every must be even, and check_every must divide interval "
                             "(retired samples keep the even-step layout)")
        lag = interval // check_every
        n = self.n_samples
        done = np.zeros(n, bool)
        steps = np.full(n, -1, np.int64)
        past = []  # <vx> per sample at the last `lag` checks
        while self._t < max_steps and not done.all():
            self.run(min(check_every, max_steps - self._t), graph_steps=graph_steps)
            r = self.reduce()
            bad = r[:, 5] > 0
            if bad.any():
                raise FloatingPointError(f"non-finite cells in samples {np.nonzero(bad)[0].tolist()} "
                                         f"after {self._t} steps")
            cur = r[:, 1] / np.maximum(r[:, 6], 1)
            if callback is not None:
                callback(self._t, r)
            if len(past) == lag:
                prev = past.pop(0)
                rel = np.abs(cur - prev) / np.maximum(np.abs(cur), 1e-300)
                newly = (~done) & (rel < tol) & (cur != 0)
                steps[newly] = self._t
                done |= newly
                if newly.any():
                    self.retire(done)
            past.append(cur)
        steps[~done] = self._t
        return steps, done
