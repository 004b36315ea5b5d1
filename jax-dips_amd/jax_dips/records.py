"""Device-resident affine residual records and the training-step engine.

A record row is one training point p of one zoom level (stencil spacing h):

    r(p) = c_p + sum_e w_e * NN_{phase_e}(fl32(p + o_e * h))

with o_e in [-2,2]^3 a lattice offset and phase_e the sign of phi at that
position (NN+ where phi >= 0, MLP.py:98).  The sum is exactly the reference's
nan_to_num(lhs/diag) - nan_to_num(rhs/diag) of compute_Ax_and_b_preconditioned_fn
(discretization.py:299-423), which is affine in the network outputs because
geometry, regression rows and callbacks are parameter-independent.

Layout in HBM (SoA, one set per zoom level):
  row_point i32[N]  grid index of the row's point      row_c  f32[N]
  row_ptr   i32[N+1] CSR offsets into the entry arrays
  ent_code  u32[nnz] offset code | phase << 7          ent_w  f32[nnz]
Rows of one batch are contiguous (the reference's contiguous batches,
data_management.py:79-160) and, inside a batch, ordered by entry count so a
wavefront sees uniform trip counts (boundary rows: 1 entry, bulk: 7,
interface: up to 125).
"""
import ctypes

import numpy as np

from . import _native as N


# rows with at most this many entries are processed one lane per row; longer
# (interface) rows one wavefront per row
BULK_MAX_ENTRIES = 8


def batch_bounds(n, batch_size):
    """Contiguous batches (DatasetDict with num_gpus = 1, batch_size clamped to n)."""
    B = min(int(batch_size), int(n))
    nb = -(-n // B)
    begins = np.arange(nb, dtype=np.int64) * B
    rows = np.minimum(begins + B, n) - begins
    return B, begins, rows


def pack_dense(W250, c, batch_size):
    """Pack dense (offset, phase)-keyed records (host arrays) into the CSR
    device layout.

    W250 (n,250) weights indexed 2*code + phase, c (n,) constants.  Returns a
    dict of host arrays in row order (batches contiguous, rows sorted by entry
    count inside each batch)."""
    n = W250.shape[0]
    B, begins, rows = batch_bounds(n, batch_size)
    nnz_row = (W250 != 0).sum(axis=1)
    order = np.empty(n, dtype=np.int64)
    for b0, nr in zip(begins, rows):
        seg = np.arange(b0, b0 + nr)
        order[b0:b0 + nr] = seg[np.argsort(nnz_row[seg], kind="stable")]
    Wo = W250[order]
    rr, cc = np.nonzero(Wo)
    row_ptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(rr, minlength=n), out=row_ptr[1:])
    if row_ptr[-1] >= 2 ** 31:
        raise ValueError("too many record entries for int32 offsets")
    code = (cc // 2).astype(np.uint32) | ((cc % 2).astype(np.uint32) << np.uint32(7))
    nz_sorted = nnz_row[order]
    nbulk = np.array([int((nz_sorted[b0:b0 + nr] <= BULK_MAX_ENTRIES).sum()) for b0, nr in zip(begins, rows)],
                     dtype=np.int64)
    return dict(row_point=order.astype(np.int32), row_ptr=row_ptr.astype(np.int32),
                row_c=np.asarray(c, dtype=np.float64)[order].astype(np.float32),
                ent_code=code, ent_w=Wo[rr, cc].astype(np.float32),
                batch_size=B, batch_begin=begins, batch_rows=rows, batch_nbulk=nbulk)


def pack_tiles(row_ptr, nrows, batch_begin, batch_rows):
    """Host packing of the rows of every batch into wavefront tiles
    (dips_pack_tiles).  Returns (tile_row int32, batch_tile_begin int64 of
    length nbatches + 1)."""
    L = N.lib()
    row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
    bb = np.ascontiguousarray(batch_begin, dtype=np.int64)
    br = np.ascontiguousarray(batch_rows, dtype=np.int64)
    btb = np.zeros(len(bb) + 1, np.int64)
    total = L.dips_pack_tiles(row_ptr.ctypes.data, int(nrows), bb.ctypes.data, br.ctypes.data, len(bb), None, 0,
                              btb.ctypes.data)
    if total < 0:
        raise N.DipsError(L.dips_last_error().decode())
    tile_row = np.zeros(int(total), np.int32)
    got = L.dips_pack_tiles(row_ptr.ctypes.data, int(nrows), bb.ctypes.data, br.ctypes.data, len(bb),
                            tile_row.ctypes.data, int(total), btb.ctypes.data)
    if got != total:
        raise N.DipsError(L.dips_last_error().decode())
    return tile_row, btb


class DeviceGrid:
    """Training points on the device: the nodes of the float32 axes ('ij' C
    order) followed by optional extra points (nextra, 3), the random points that
    pad a short last batch.  Empty axes and only extra points: a scattered
    training set (from_points)."""

    @classmethod
    def from_points(cls, R, device="cuda"):
        """A set of scattered points (N, 3), no grid nodes: the training points
        of a gstate built by mesh.init_gstate_3d_manually (reference
        domain/mesh.py:179-213)."""
        empty = np.zeros(0, np.float32)
        return cls(empty, empty, empty, extra=np.asarray(R, np.float32).reshape(-1, 3), device=device)

    def __init__(self, xs, ys, zs, extra=None, device="cuda"):
        import torch

        self.xs = torch.as_tensor(np.asarray(xs, np.float32), device=device)
        self.ys = torch.as_tensor(np.asarray(ys, np.float32), device=device)
        self.zs = torch.as_tensor(np.asarray(zs, np.float32), device=device)
        self.shape = (len(xs), len(ys), len(zs))
        self.ngrid = len(xs) * len(ys) * len(zs)
        ex = np.zeros((0, 3), np.float32) if extra is None else np.asarray(extra, np.float32).reshape(-1, 3)
        self.extra = torch.as_tensor(np.ascontiguousarray(ex), device=device) if len(ex) else None
        self.nextra = int(len(ex))
        self.npts = self.ngrid + self.nextra
        if self.ngrid == 0 and self.nextra == 0:
            raise ValueError("DeviceGrid: no points")
        nodes = self.ngrid > 0
        self.c = N.Grid(self.xs.data_ptr() if nodes else None, self.ys.data_ptr() if nodes else None,
                        self.zs.data_ptr() if nodes else None, *(self.shape if nodes else (0, 0, 0)),
                        self.extra.data_ptr() if self.extra is not None else None, self.nextra)

    def points(self):
        """(npts, 3) float32 device tensor of every training point."""
        import torch

        X, Y, Z = torch.meshgrid(self.xs, self.ys, self.zs, indexing="ij")
        P = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], dim=1)
        return P if self.extra is None else torch.cat([P, self.extra], dim=0)


class DeviceRecords:
    """Records of one zoom level in device memory, plus the batch table."""

    def __init__(self, packed, h, device="cuda"):
        import torch

        def T(a, dt):
            return torch.as_tensor(np.ascontiguousarray(a), device=device).to(dt)

        self.row_point = T(packed["row_point"], torch.int32)
        self.row_ptr = T(packed["row_ptr"], torch.int32)
        self.row_c = T(packed["row_c"], torch.float32)
        nnz = len(packed["ent_w"])
        codes = packed["ent_code"].astype(np.int64).astype(np.int32) if nnz else np.zeros(1, np.int32)
        self.ent_code = T(codes, torch.int32)
        self.ent_w = T(packed["ent_w"] if nnz else np.zeros(1, np.float32), torch.float32)
        self.nrows = len(packed["row_c"])
        self.nnz = nnz
        self.h = np.asarray(h, np.float32)
        self.batch_size = int(packed["batch_size"])
        self.batch_begin = np.asarray(packed["batch_begin"], np.int64)
        self.batch_rows = np.asarray(packed["batch_rows"], np.int64)
        self.batch_nbulk = np.asarray(packed["batch_nbulk"], np.int64)
        self.row_coeffs = None
        self._finish()

    @classmethod
    def from_device(cls, row_point, row_ptr, row_c, ent_code, ent_w, nnz, h, batch_size, batch_begin, batch_rows,
                    batch_nbulk, row_coeffs=None):
        """Wrap records already in device memory (built by the geometry kernels)."""
        self = cls.__new__(cls)
        self.row_point, self.row_ptr, self.row_c = row_point, row_ptr, row_c
        self.ent_code, self.ent_w = ent_code, ent_w
        self.nrows = int(row_c.numel())
        self.nnz = int(nnz)
        self.h = np.asarray(h, np.float32)
        self.batch_size = int(batch_size)
        self.batch_begin = np.asarray(batch_begin, np.int64)
        self.batch_rows = np.asarray(batch_rows, np.int64)
        self.batch_nbulk = np.asarray(batch_nbulk, np.int64)
        self.row_coeffs = row_coeffs
        self._finish()
        return self

    def _finish(self):
        self.c = N.Records(self.row_point.data_ptr(), self.row_ptr.data_ptr(), self.row_c.data_ptr(),
                           self.ent_code.data_ptr(), self.ent_w.data_ptr(), self.nrows, self.nnz,
                           float(self.h[0]), float(self.h[1]), float(self.h[2]),
                           self.row_coeffs.data_ptr() if self.row_coeffs is not None else None)

    def set_point_coeffs(self, coeffs):
        """Attach the 26 cell coefficients of every point ((N,26), point order)
        as the per-row preconditioner input."""
        import torch

        C = torch.as_tensor(np.asarray(coeffs), device=self.row_point.device).to(torch.float32)
        if C.shape != (self.nrows, 26):
            raise ValueError(f"expected ({self.nrows}, 26) coefficients, got {tuple(C.shape)}")
        self.row_coeffs = C[self.row_point.to(torch.int64)].contiguous()
        self._finish()

    def mfma_batches(self, grid=None, table_order=False):
        """Phase-sorted entry arrays (+ entries first inside each batch) and one
        dips_mfma_batch_t per batch, for the MFMA and hashgrid kernels (built
        once per ordering, cached).  grid (DeviceGrid): the entries of each
        (batch, phase) segment are also ordered by their position fl32(p + o h),
        the position the kernels compute, so that entries at one position (a
        stencil neighbour of one point that rounds onto the next point) are
        adjacent; the hashgrid backward then sends one set of table-gradient
        records for them, and the batches carry these runs as groups
        (t_group, g_begin/g_plus/g_minus) whose network the MFMA kernels
        evaluate once.  Any order inside a segment is valid: t_csr maps each
        entry to its CSR slot.  The position order is lexicographic, x-major
        with z fastest (the order of the grid points, hence of the rows, which
        keeps the MFMA row sums' gathers of u per item local), or with
        table_order z-major with x fastest (the order of the hashgrid's dense
        table rows, x + y res + z res^2, for the hashgrid kernels' table
        gathers and record aggregation)."""
        import torch

        cache = "_mfma" if grid is None else ("_mfma_tab" if table_order else "_mfma_pos")
        if getattr(self, cache, None) is not None:
            return getattr(self, cache)
        dev = self.row_ptr.device
        ptr = self.row_ptr.to(torch.int64)
        counts = ptr[1:] - ptr[:-1]
        n = self.nnz
        ent_row = torch.repeat_interleave(torch.arange(self.nrows, device=dev, dtype=torch.int64), counts)
        code = self.ent_code[:n].to(torch.int64)
        phase = (code >> 7) & 1
        batch = ent_row // self.batch_size
        key = batch * 2 + (1 - phase)
        if grid is None:
            _, perm = torch.sort(key, stable=True)
        else:
            # offsets o in [-2, 2]^3 from bits 0..6 of the code (25 (oz+2) + 5 (oy+2) + (ox+2))
            c = code & 127
            o = torch.stack([c % 5 - 2, (c // 5) % 5 - 2, c // 25 - 2], dim=1).to(torch.float32)
            h = torch.as_tensor(self.h, dtype=torch.float32, device=dev)
            # o h is exact for |o| <= 2, so the sum rounds once, as the kernels' fused form does
            X = grid.points()[self.row_point.to(torch.int64)[ent_row]] + o * h
            del o, c
            perm = torch.arange(n, device=dev)
            # least significant key first, then the segment: z, y, x (x-major, z fastest, as the rows are) or,
            # in table order, x, y, z (z-major, x fastest, as the hashgrid's dense table rows are)
            for col in ((0, 1, 2) if table_order else (2, 1, 0)):
                _, p2 = torch.sort(X[perm, col], stable=True)
                perm = perm[p2]
            _, p2 = torch.sort(key[perm], stable=True)
            perm = perm[p2]
            # groups: runs of bit-identical positions inside a (batch, phase) segment of this order
            Xs = X[perm]
            del X
            ks = key[perm]
            new = torch.ones(n, dtype=torch.bool, device=dev)
            if n > 1:
                new[1:] = (ks[1:] != ks[:-1]) | (Xs[1:] != Xs[:-1]).any(dim=1)
            del Xs, ks
            starts = torch.nonzero(new).flatten()
            t_group = torch.cat([starts, torch.tensor([n], device=dev, dtype=torch.int64)]).to(torch.int32).contiguous()
        t_row = ent_row[perm].to(torch.int32).contiguous()
        t_code = self.ent_code[:n][perm].contiguous()
        t_w = self.ent_w[:n][perm].contiguous()
        t_csr = perm.to(torch.int32).contiguous()
        nplus = torch.zeros(len(self.batch_begin), dtype=torch.int64, device=dev).index_add_(0, batch, phase).cpu().numpy()
        ptr_h = ptr.cpu().numpy()
        if grid is not None:
            # each batch's groups: + groups from its first entry, - groups from its first - entry
            bnd = np.array([[int(ptr_h[b0]), int(ptr_h[b0]) + int(nplus[b]), int(ptr_h[b0 + nr])]
                            for b, (b0, nr) in enumerate(zip(self.batch_begin, self.batch_rows))], np.int64)
            gidx = torch.searchsorted(starts, torch.as_tensor(bnd.ravel(), device=dev)).cpu().numpy().reshape(-1, 3)
        batches = []
        for b, (b0, nr) in enumerate(zip(self.batch_begin, self.batch_rows)):
            e0, e1 = int(ptr_h[b0]), int(ptr_h[b0 + nr])
            bt = N.MfmaBatch(int(b0), int(nr), e0, e1, t_row.data_ptr(), t_code.data_ptr(), t_w.data_ptr(),
                             t_csr.data_ptr(), e0, int(nplus[b]), int(e1 - e0 - nplus[b]), int(e1 - e0))
            if grid is not None:
                bt.t_group = t_group.data_ptr()
                bt.g_begin, bt.g_plus, bt.g_minus = (int(gidx[b, 0]), int(gidx[b, 1] - gidx[b, 0]),
                                                     int(gidx[b, 2] - gidx[b, 1]))
            batches.append(bt)
        mb = dict(t_row=t_row, t_code=t_code, t_w=t_w, t_csr=t_csr, batches=batches,
                  max_nnz=int(max(int(ptr_h[b0 + nr] - ptr_h[b0]) for b0, nr in
                                  zip(self.batch_begin, self.batch_rows))))
        if grid is not None:
            mb["t_group"] = t_group
            mb["groups"] = int(starts.numel())
        setattr(self, cache, mb)
        return mb

    def mfma_items(self, grid, mb):
        """Item positions and the entry -> item map of the batches in mb (one
        ordering of mfma_batches), which the MFMA step kernels read: built on
        the device once per ordering (dips_mfma_items per batch) and attached
        to every batch (i_pos, i_count, e_item).  Items are the groups when mb
        has them, else the entries."""
        import torch

        if "i_pos" in mb:
            return mb
        dev = self.row_ptr.device
        ni = mb["groups"] if "t_group" in mb else self.nnz
        i_pos = torch.zeros(3 * max(ni, 1), dtype=torch.float32, device=dev)
        e_item = torch.zeros(max(self.nnz, 1), dtype=torch.int32, device=dev)
        for bt in mb["batches"]:
            bt.i_count = ni
            N.check(N.lib().dips_mfma_items(ctypes.byref(grid.c), ctypes.byref(self.c), ctypes.byref(bt), N.ptr(i_pos),
                                            N.ptr(e_item), N.current_stream()), "dips_mfma_items")
            bt.i_pos, bt.e_item = i_pos.data_ptr(), e_item.data_ptr()
        mb["i_pos"], mb["e_item"] = i_pos, e_item
        return mb

    def tiles(self, grid):
        """Wavefront tiles of every batch (rows packed into tiles of at most 64
        slots, dips_pack_tiles) and their slot layout (dips_tile_slots), built
        once per level and cached."""
        import torch

        if getattr(self, "_tiles", None) is not None:
            return self._tiles
        L = N.lib()
        ptr_h = np.ascontiguousarray(self.row_ptr.cpu().numpy(), dtype=np.int32)
        tile_row, btb = pack_tiles(ptr_h, self.nrows, self.batch_begin, self.batch_rows)
        dev = self.row_ptr.device
        tr = torch.as_tensor(tile_row, device=dev)
        nb = len(tile_row)
        slots = {k: torch.empty(nb * 64, dtype=torch.float32, device=dev) for k in ("x", "y", "z", "w", "c")}
        slots["info"] = torch.empty(nb * 64, dtype=torch.int32, device=dev)
        slots["row"] = torch.empty(nb * 64, dtype=torch.int32, device=dev)
        keys = ("x", "y", "z", "w", "c", "info", "row")
        N.check(L.dips_tile_slots(ctypes.byref(grid.c), ctypes.byref(self.c), N.ptr(tr), nb,
                                  *(N.ptr(slots[k]) for k in keys), N.current_stream()), "dips_tile_slots")
        c = N.Tiles(tr.data_ptr(), nb, *(slots[k].data_ptr() for k in keys))
        self._tiles = dict(tile_row=tr, batch_tile_begin=btb, slots=slots, c=c)
        return self._tiles

    def to_dense(self):
        """Host copy as dense (N,125) weights, (N,125) phase and (N,) constants in
        grid-point order (for tests and inspection)."""
        rp = self.row_point.cpu().numpy().astype(np.int64)
        ptr = self.row_ptr.cpu().numpy().astype(np.int64)
        code = self.ent_code.cpu().numpy().astype(np.int64)[: self.nnz]
        w = self.ent_w.cpu().numpy()[: self.nnz]
        c = self.row_c.cpu().numpy()
        n = self.nrows
        W = np.zeros((n, 250), np.float64)
        rows = np.repeat(np.arange(n), np.diff(ptr))
        slot = (code & 127) * 2 + ((code >> 7) & 1)
        np.add.at(W, (rp[rows], slot), w)
        C = np.zeros(n, np.float64)
        C[rp] = c
        return W, C


OPT_KIND = {"custom": 0, "adam": 1, "rmsprop": 2}
SCHED_KIND = {"exponential": 0, "polynomial": 1}


def make_opt(optimizer_dict):
    name = optimizer_dict.get("optimizer_name", "custom")
    if name not in OPT_KIND:
        raise ValueError(f"Unknown optimizer: {name}")
    sched = optimizer_dict.get("sched", {}) or {}
    sname = sched.get("scheduler_name", "exponential")
    if sname not in SCHED_KIND:
        raise ValueError(f"Unknown scheduler: {sname}")
    return N.Opt(OPT_KIND[name], SCHED_KIND[sname], float(optimizer_dict.get("learning_rate", 1e-3)),
                 float(sched.get("decay_rate", 0.9)), 1000.0, 1.0)


class Engine:
    """Parameters, optimizer state and scratch of one two-phase network on the GPU."""

    def __init__(self, dims, params, optimizer_dict, max_rows, max_steps, device="cuda", precond=None, act=0):
        """precond: None, or {"widths": [w1] or [w1, w2], "scale": s} for the
        learned preconditioner; its parameters follow the two networks' in
        params, and records then need per-row cell coefficients.  act: hidden
        activation, 0 = tanh, 1 = CELU (tile kernel only)."""
        import torch

        self.dims = tuple(int(d) for d in dims)
        self.precond = None
        pc = (0, 0, 0.0)
        if precond is not None:
            widths = [int(w) for w in precond["widths"]]
            if not 1 <= len(widths) <= 2 or min(widths) < 1:
                raise ValueError(f"preconditioner widths must be one or two positive widths, got {widths}")
            self.precond = dict(widths=widths, scale=float(precond.get("scale", 1.0)))
            pc = (widths[0], widths[1] if len(widths) == 2 else 0, self.precond["scale"])
        self.act = int(act)
        self.net = N.Net(*self.dims, *pc, self.act)
        L = N.lib()
        self.lane_path = bool(L.dips_net_supported(ctypes.byref(self.net)))
        self.tile_path = bool(L.dips_tile_supported(ctypes.byref(self.net)))
        self.mfma_path = bool(L.dips_mfma_supported(ctypes.byref(self.net)))
        if (self.precond is not None or self.act != 0) and not self.tile_path:
            raise N.DipsError(f"network shape {self.dims} (activation {self.act}, preconditioner "
                              f"{self.precond}) has no compiled tile kernel")
        if not (self.lane_path or self.mfma_path or self.tile_path):
            raise N.DipsError(f"network shape (Lm,Hm,Lp,Hp)={self.dims} has no compiled kernel")
        self.P = int(L.dips_param_count(ctypes.byref(self.net)))
        params = np.asarray(params, np.float32)
        if params.size != self.P:
            raise ValueError(f"expected {self.P} parameters, got {params.size}")
        self.device = device
        self.params = torch.as_tensor(params, device=device).clone()
        self.m = torch.zeros(self.P, dtype=torch.float32, device=device)
        self.v = torch.zeros(self.P, dtype=torch.float32, device=device)
        self.count = torch.zeros(1, dtype=torch.int32, device=device)
        self.loss_hist = torch.zeros(max(1, int(max_steps)), dtype=torch.float32, device=device)
        # per-workgroup partial rows of the lane-per-row kernel (the MFMA path keeps its own workspace)
        self.partial_cap = int(L.dips_partial_size(ctypes.byref(self.net), int(max_rows))) if self.lane_path else 1
        if self.tile_path:
            self.partial_cap = max(self.partial_cap, int(L.dips_tile_partial_size(ctypes.byref(self.net))))
        self.partial = torch.zeros(self.partial_cap, dtype=torch.float32, device=device)
        self.sync = torch.zeros(1, dtype=torch.int32, device=device)  # arrival counter of the fused step
        self.fused = True
        self.use_tiles = True  # packed-tile kernel when the shape has one, else the per-row kernel
        # tile steps as two launches (step kernel, then a one-workgroup reduction + optimizer kernel),
        # which is faster than the single launch whose last workgroup applies the optimizer
        self.tile_single_launch = False
        # MFMA path: one network evaluation per group of entries at one position (position-ordered batches)
        self.mfma_grouped = True
        self.opt = make_opt(optimizer_dict)

    def loss_grad(self, grid, rec, row_begin, nrows, inv_batch, want_grad=True, nbulk=None):
        """(loss, grad) of rows [row_begin, row_begin+nrows) as float64 numpy.
        nbulk defaults to the batch table's split when the range is a batch."""
        import torch

        if nbulk is None:
            hit = np.nonzero((rec.batch_begin == row_begin) & (rec.batch_rows == nrows))[0]
            nbulk = int(rec.batch_nbulk[hit[0]]) if hit.size else 0
        nblk = ctypes.c_int32(0)
        st = N.lib().dips_loss_grad(ctypes.byref(grid.c), ctypes.byref(rec.c), int(row_begin), int(nrows), int(nbulk),
                                    ctypes.byref(self.net), N.ptr(self.params), N.ptr(self.partial),
                                    self.partial_cap, int(bool(want_grad)), ctypes.byref(nblk), N.current_stream())
        N.check(st, "dips_loss_grad")
        torch.cuda.synchronize()
        part = self.partial[: nblk.value * (self.P + 1)].view(nblk.value, self.P + 1).double().cpu().numpy()
        s = part.sum(axis=0) * float(inv_batch)
        return s[self.P], s[: self.P]

    def tile_loss_grad(self, grid, rec, b, inv_batch, want_grad=True):
        """(loss, grad) of batch b through the tile kernel, float64 numpy."""
        import torch

        tl = rec.tiles(grid)
        tb = int(tl["batch_tile_begin"][b])
        nt = int(tl["batch_tile_begin"][b + 1]) - tb - 1
        nblk = ctypes.c_int32(0)
        N.check(N.lib().dips_tile_loss_grad(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(tl["c"]), tb, nt,
                                            ctypes.byref(self.net), N.ptr(self.params), N.ptr(self.partial),
                                            self.partial_cap, int(bool(want_grad)), ctypes.byref(nblk),
                                            N.current_stream()), "dips_tile_loss_grad")
        torch.cuda.synchronize()
        part = self.partial[: nblk.value * (self.P + 1)].view(nblk.value, self.P + 1).double().cpu().numpy()
        s = part.sum(axis=0) * float(inv_batch)
        return s[self.P], s[: self.P]

    def train_level_tiles(self, grid, rec, nepochs):
        tl = rec.tiles(grid)
        nb = len(rec.batch_begin)
        st = N.lib().dips_tile_train_level(
            ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(tl["c"]), N.i64_array(tl["batch_tile_begin"]), nb,
            rec.batch_size, int(nepochs), ctypes.byref(self.net), ctypes.byref(self.opt), N.ptr(self.params),
            N.ptr(self.m), N.ptr(self.v), N.ptr(self.count), N.ptr(self.loss_hist), int(self.loss_hist.numel()),
            N.ptr(self.partial), self.partial_cap, N.ptr(self.sync) if self.tile_single_launch else None,
            N.current_stream())
        N.check(st, "dips_tile_train_level")

    def mfma_workspace(self, rec, grid, grouped=True):
        """Phase-sorted batches of rec for the MFMA kernels (grouped: in
        position order, with groups of entries at one position), their items,
        and a workspace large enough for them."""
        import torch

        mb = rec.mfma_items(grid, rec.mfma_batches(grid if grouped else None))
        need = int(N.lib().dips_mfma_workspace(ctypes.byref(self.net), int(rec.batch_size), mb["max_nnz"]))
        if getattr(self, "_mwork", None) is None or self._mwork.numel() < need:
            self._mwork = torch.empty(need, dtype=torch.float32, device=self.device)
            self._mdwork = torch.empty(int(rec.batch_size) // 256 + self.P // 64 + 64, dtype=torch.float64,
                                       device=self.device)
        return mb

    def mfma_loss_grad(self, grid, rec, b, inv_batch, grouped=True):
        """(loss, grad) of batch b through the MFMA kernels, float64 numpy;
        grouped: entries at one position evaluated once (else each entry)."""
        import torch

        mb = self.mfma_workspace(rec, grid, grouped)
        g = torch.empty(self.P, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float64, device=self.device)
        N.check(N.lib().dips_mfma_loss_grad(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(mb["batches"][b]),
                                            ctypes.byref(self.net), ctypes.c_float(inv_batch), N.ptr(self.params),
                                            N.ptr(g), N.ptr(loss), N.ptr(self._mwork), int(self._mwork.numel()),
                                            N.ptr(self._mdwork), int(self._mdwork.numel()), N.current_stream()),
                "dips_mfma_loss_grad")
        torch.cuda.synchronize()
        return float(loss.item()), g.double().cpu().numpy()

    def train_level_mfma(self, grid, rec, nepochs):
        mb = self.mfma_workspace(rec, grid, self.mfma_grouped)
        L = N.lib()
        inv_b = ctypes.c_float(1.0 / rec.batch_size)
        for _ in range(int(nepochs)):
            for bt in mb["batches"]:
                N.check(L.dips_mfma_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(bt),
                                               ctypes.byref(self.net), ctypes.byref(self.opt), inv_b,
                                               N.ptr(self.params), N.ptr(self.m), N.ptr(self.v), N.ptr(self.count),
                                               N.ptr(self.loss_hist), int(self.loss_hist.numel()), N.ptr(self._mwork),
                                               int(self._mwork.numel()), N.ptr(self._mdwork),
                                               int(self._mdwork.numel()), N.current_stream()), "dips_mfma_train_step")

    def train_level(self, grid, rec, nepochs):
        if self.tile_path and (self.use_tiles or self.precond is not None or self.act != 0):
            return self.train_level_tiles(grid, rec, nepochs)
        if not self.lane_path:
            return self.train_level_mfma(grid, rec, nepochs)
        st = N.lib().dips_train_level(
            ctypes.byref(grid.c), ctypes.byref(rec.c),
            N.i64_array(rec.batch_begin), N.i64_array(rec.batch_rows), N.i64_array(rec.batch_nbulk),
            len(rec.batch_begin), rec.batch_size, int(nepochs), ctypes.byref(self.net), ctypes.byref(self.opt),
            N.ptr(self.params), N.ptr(self.m), N.ptr(self.v), N.ptr(self.count), N.ptr(self.loss_hist),
            int(self.loss_hist.numel()), N.ptr(self.partial), self.partial_cap,
            N.ptr(self.sync) if self.fused else None, N.current_stream())
        N.check(st, "dips_train_level")

    def prepare(self, grid, rec):
        """Build what train_level needs for rec (tiles, MFMA batches and
        workspace) ahead of time, outside any graph capture."""
        if self.tile_path and (self.use_tiles or self.precond is not None or self.act != 0):
            rec.tiles(grid)
        elif not self.lane_path:
            self.mfma_workspace(rec, grid, self.mfma_grouped)

    def capture_epoch(self, grid, rec):
        """One epoch over rec's batches (the launches of train_level(grid, rec, 1))
        captured in a HIP graph; each replay() trains one more epoch (step
        counter, schedule and loss history advance on the device)."""
        return capture_epoch(self, grid, rec)


def capture_epoch(engine, grid, rec):
    """HIP graph of engine.train_level(grid, rec, 1): every launch of an epoch
    goes to the capture stream; nothing runs until replay()."""
    import torch

    engine.prepare(grid, rec)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        engine.train_level(grid, rec, 1)
    return g


class HashEngine:
    """Parameters, optimizer state and workspace of the hashgrid network on the
    GPU (dips_hash_*): the two CELU MLPs and the encoder's table, trained over
    the phase-sorted entries of each batch, ordered by position inside each
    phase (DeviceRecords.mfma_batches with the grid)."""

    def __init__(self, hashgrid, params, optimizer_dict, max_steps, device="cuda"):
        """hashgrid: (levels, features, table_size, n_min, n_max, bound, encoder)
        of dips_hashgrid_t; params: flat float32 (sol_p, sol_m, then the table
        of the hashgrid encoder)."""
        import torch

        self.hg = N.HashGrid(*[t(v) for t, v in zip((int, int, int, int, int, float, int), hashgrid)])
        L = N.lib()
        self.P = int(L.dips_hash_param_count(ctypes.byref(self.hg)))
        if self.P < 0:
            raise N.DipsError(f"hashgrid {tuple(hashgrid)}: {N.lib().dips_last_error().decode()}")
        params = np.asarray(params, np.float32)
        if params.size != self.P:
            raise ValueError(f"expected {self.P} parameters, got {params.size}")
        self.device = device
        self.params = torch.as_tensor(params, device=device).clone()
        self.m = torch.zeros(self.P, dtype=torch.float32, device=device)
        self.v = torch.zeros(self.P, dtype=torch.float32, device=device)
        self.count = torch.zeros(1, dtype=torch.int32, device=device)
        self.loss_hist = torch.zeros(max(1, int(max_steps)), dtype=torch.float32, device=device)
        self.opt = make_opt(optimizer_dict)
        self._work = None

    def workspace(self, rec, grid):
        """Entry arrays of rec ordered by position inside each (batch, phase)
        segment, and a workspace large enough for its batches."""
        import torch

        mb = rec.mfma_batches(grid, table_order=True)
        dbl = ctypes.c_int64(0)
        need = int(N.lib().dips_hash_workspace(ctypes.byref(self.hg), int(rec.batch_size), mb["max_nnz"],
                                               ctypes.byref(dbl)))
        if need < 0:
            raise N.DipsError(N.lib().dips_last_error().decode())
        if self._work is None or self._work.numel() < need or self._dwork.numel() < dbl.value:
            self._work = torch.empty(need, dtype=torch.float32, device=self.device)
            self._dwork = torch.empty(dbl.value, dtype=torch.float64, device=self.device)
        return mb

    def loss_grad(self, grid, rec, b, inv_batch):
        """(loss, grad) of batch b, float64 numpy."""
        import torch

        mb = self.workspace(rec, grid)
        g = torch.empty(self.P, dtype=torch.float32, device=self.device)
        loss = torch.empty(1, dtype=torch.float64, device=self.device)
        N.check(N.lib().dips_hash_loss_grad(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(mb["batches"][b]),
                                            ctypes.byref(self.hg), ctypes.c_float(inv_batch), N.ptr(self.params),
                                            N.ptr(g), N.ptr(loss), N.ptr(self._work), int(self._work.numel()),
                                            N.ptr(self._dwork), int(self._dwork.numel()), N.current_stream()),
                "dips_hash_loss_grad")
        torch.cuda.synchronize()
        return float(loss.item()), g.double().cpu().numpy()

    def train_level(self, grid, rec, nepochs):
        mb = self.workspace(rec, grid)
        L = N.lib()
        inv_b = ctypes.c_float(1.0 / rec.batch_size)
        for _ in range(int(nepochs)):
            for bt in mb["batches"]:
                N.check(L.dips_hash_train_step(ctypes.byref(grid.c), ctypes.byref(rec.c), ctypes.byref(bt),
                                               ctypes.byref(self.hg), ctypes.byref(self.opt), inv_b,
                                               N.ptr(self.params), N.ptr(self.m), N.ptr(self.v), N.ptr(self.count),
                                               N.ptr(self.loss_hist), int(self.loss_hist.numel()), N.ptr(self._work),
                                               int(self._work.numel()), N.ptr(self._dwork),
                                               int(self._dwork.numel()), N.current_stream()), "dips_hash_train_step")


class BatchEngine:
    """Parameters and optimizer state of many independent problems with one
    network shape, trained together (dips_batch_setup / dips_batch_steps: per
    global step one persistent launch whose workgroups take (problem, block)
    items from a device queue, then one optimizer workgroup per problem).  With
    the default shape each problem follows the trajectory Engine.train_level
    gives it alone; autotune() picks a faster shape for a set of problems."""

    def __init__(self, dims, params_list, optimizer_dict, max_steps, device="cuda", precond=None, act=0):
        """max_steps: one optimizer-step capacity (loss history length) per problem."""
        self.engines = [Engine(dims, p, optimizer_dict, max_rows=1, max_steps=ms, device=device, precond=precond,
                               act=act) for p, ms in zip(params_list, max_steps)]
        if len(self.engines) != len(params_list):
            raise ValueError("one max_steps entry per problem")
        if not self.engines[0].tile_path:
            raise N.DipsError(f"network shape {tuple(dims)} has no compiled tile kernel for batched training")
        self.net = self.engines[0].net
        self.opt = self.engines[0].opt
        self.device = device
        self._work = None

    def problems(self, grids, recs):
        """The dips_batch_problem_t table (host) for these grids and records."""
        n = len(self.engines)
        if len(grids) != n or len(recs) != n:
            raise ValueError("one grid and one record set per problem")
        probs = (N.BatchProblem * n)()
        keep = []
        for i, (eng, g, rec) in enumerate(zip(self.engines, grids, recs)):
            tl = rec.tiles(g)
            btb = N.i64_array(tl["batch_tile_begin"])
            keep.append(btb)
            q = probs[i]
            q.grid, q.rec, q.tiles = g.c, rec.c, tl["c"]
            q.batch_tile_begin = ctypes.cast(btb, ctypes.c_void_p)
            q.nbatches = len(rec.batch_begin)
            q.batch_size = int(rec.batch_size)
            q.params, q.m, q.v = eng.params.data_ptr(), eng.m.data_ptr(), eng.v.data_ptr()
            q.count, q.loss_hist = eng.count.data_ptr(), eng.loss_hist.data_ptr()
            q.loss_cap = int(eng.loss_hist.numel())
        return probs, keep

    def setup(self, grids, recs, nepochs, cfg=None):
        """dips_batch_setup for nepochs epochs of these records (copies the
        problem table, synchronizes); cfg = (blk_cap, grid_blocks) or None for
        the default, which reproduces the single-problem steps bit for bit.
        Returns the plan for steps()."""
        import torch

        L = N.lib()
        probs, keep = self.problems(grids, recs)
        n = len(self.engines)
        c = N.BatchCfg(*(cfg if cfg is not None else (0, 0)))
        need = int(L.dips_batch_workspace_cfg(probs, n, ctypes.byref(self.net), ctypes.byref(c)))
        if need < 0:
            N.check(-1, "dips_batch_workspace_cfg")
        if self._work is None or self._work.numel() < need:
            self._work = torch.empty(need, dtype=torch.uint8, device=self.device)
        plan = N.BatchPlan()
        N.check(L.dips_batch_setup(probs, n, int(nepochs), ctypes.byref(self.net), ctypes.byref(self.opt),
                                   ctypes.byref(c), N.ptr(self._work), int(self._work.numel()), ctypes.byref(plan),
                                   N.current_stream()), "dips_batch_setup")
        del keep
        return plan

    def steps(self, plan, nsteps):
        """Enqueue nsteps global steps (launches only)."""
        N.check(N.lib().dips_batch_steps(ctypes.byref(self.net), ctypes.byref(plan), int(nsteps), N.ptr(self._work),
                                         N.current_stream()), "dips_batch_steps")

    def train_level(self, grids, recs, nepochs, cfg=None, graph=True):
        """nepochs epochs of every problem at one zoom level: one setup, then
        the global steps of an epoch (the most batches over the problems)
        captured once in a HIP graph and replayed per epoch (graph=False:
        launched directly)."""
        import torch

        plan = self.setup(grids, recs, nepochs, cfg)
        if nepochs <= 0:
            return plan
        per_epoch = plan.max_steps // int(nepochs)
        if graph and nepochs > 1:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self.steps(plan, per_epoch)
            for _ in range(int(nepochs)):
                g.replay()
            del g
        else:
            self.steps(plan, plan.max_steps)
        return plan

    def _snapshot(self):
        return [[t.clone() for t in (e.params, e.m, e.v, e.count, e.loss_hist)] for e in self.engines]

    def _restore(self, snap):
        for e, s in zip(self.engines, snap):
            for t, v in zip((e.params, e.m, e.v, e.count, e.loss_hist), s):
                t.copy_(v)

    def autotune(self, grids, recs, caps=(256, 128, 64, 32, 16), steps=6, cache_path=None):
        """Choose the batched step's shape (dips_batch_cfg_t: blocks per
        problem and step, persistent grid size) by timing `steps` global steps
        of these records for every candidate, from the same state (restored
        after each).  Results are cached in memory and in a JSON file keyed by
        (GPU architecture, network shape, the problems' point counts and batch
        sizes); returns the chosen (blk_cap, grid_blocks)."""
        import json
        import os

        import torch

        arch = torch.cuda.get_device_properties(torch.cuda.current_device()).gcnArchName
        sizes = sorted((int(g.npts), int(r.batch_size)) for g, r in zip(grids, recs))
        key = json.dumps([arch, [int(self.net.Lm), int(self.net.Hm), int(self.net.Lp), int(self.net.Hp),
                                 int(self.net.pc_w1), int(self.net.pc_w2), int(self.net.act)], sizes])
        path = cache_path or os.environ.get("DIPS_AUTOTUNE_CACHE",
                                            os.path.join(os.path.expanduser("~"), ".cache", "jax_dips_amd",
                                                         "batch_autotune.json"))
        if key in _BATCH_TUNED:
            return _BATCH_TUNED[key]
        try:
            with open(path) as f:
                disk = json.load(f)
        except (OSError, ValueError):
            disk = {}
        if key in disk:
            _BATCH_TUNED[key] = tuple(disk[key]["cfg"])
            return _BATCH_TUNED[key]
        snap = self._snapshot()
        # persistent grids of one and two workgroups per resident slot
        resident = int(self.setup(grids, recs, 1, (0, 0)).grid_blocks)
        timings = {}
        for cap in caps:
            for grid in (resident, 2 * resident):
                plan = self.setup(grids, recs, 1, (cap, grid))
                self.steps(plan, 1)  # warm-up
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record()
                self.steps(plan, steps)
                ev[1].record()
                torch.cuda.synchronize()
                timings[(cap, grid)] = ev[0].elapsed_time(ev[1]) / steps
                self._restore(snap)
        torch.cuda.synchronize()
        best = min(timings, key=timings.get)
        self.tuning = {"candidates_ms": {f"{c}/{g}": t for (c, g), t in timings.items()}, "cfg": list(best),
                       "resident_blocks": resident}
        _BATCH_TUNED[key] = best
        disk[key] = {"cfg": list(best), "ms_per_step": timings[best]}
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            with open(path, "w") as f:
                json.dump(disk, f)
        except OSError:
            pass
        return best


# in-process cache of BatchEngine.autotune: key -> (blk_cap, grid_blocks)
_BATCH_TUNED = {}
