"""Device-side orchestration: sequence tables in HBM, launch shapes and the
calls into libgkmhip (include/gkmhip.h) that every public function uses.

Layout on the GPU (torch tensors, caller-owned device memory):
  DeviceSeqs.planes / planes_rc  [n, nw, 2] int32  bit-planes of each sequence
                                                   and of its reverse complement
  DeviceSeqs.codes               [n, Lmax] uint8   base codes
  hist                           [B, S, nm] int64  mismatch histograms (uint64 bits)
SV tables (SVModel) are sorted by length so that each 64-SV group (one wave's
lanes) holds SVs of nearly one length, and stay resident until dropped.
"""

from __future__ import annotations

import contextlib

import math

import numpy as np
import torch

from . import _lib
from .kernels import DirectGkmKernel, GkmKernel

WAVE = 64
# Histogram bytes per cross-kernel call (rows are chunked to stay under it)
MAX_HIST_BYTES = 1 << 30


GPU_NAMES = ("rocm", "cuda", "gpu", "auto")


def placement(device) -> torch.device | None:
    """Where a model created or loaded with `device` is held: the visible
    ROCm GPU for "rocm", "cuda", "gpu" or "cuda:N" (or a torch.device);
    None -- held on the host, where it cannot compute until rocm() -- for
    "cpu", and for "auto" on a machine without a ROCm GPU ("auto" takes the
    GPU when one is visible, as the reference's backend.py:150-164 takes
    CUDA).  ValueError for any other name."""
    name = "rocm" if device is None else str(device)
    if name == "cpu" or (name == "auto" and not torch.cuda.is_available()):
        return None
    if name in GPU_NAMES or name.startswith("cuda:"):
        return device_of(None if name in GPU_NAMES else device)
    raise ValueError(f"unknown device {device!r}: use 'rocm' (or 'cuda', 'gpu', 'cuda:N'), "
                     "'auto' or 'cpu'")


def device_of(device=None) -> torch.device:
    if device is None or str(device) in GPU_NAMES:
        if not torch.cuda.is_available():
            raise RuntimeError(
                "no ROCm GPU visible: the gkm engine runs on MI355X (gfx950) only"
            )
        return torch.device("cuda", torch.cuda.current_device())
    d = torch.device(device)
    if d.type != "cuda":
        raise RuntimeError(f"the gkm engine computes on ROCm GPUs only, not {device!r}")
    return d


def stream_handle(stream=None):
    s = torch.cuda.current_stream() if stream is None else stream
    return s.cuda_stream or None


def _p(t):
    return None if t is None else t.data_ptr()


class DeviceSeqs:
    """Bit-planes (and, on request, packed windows) of sequences on the GPU."""

    def __init__(self, codes: np.ndarray, lengths: np.ndarray, device, *, rc: bool = True,
                 stream=None):
        codes = np.ascontiguousarray(codes, dtype=np.uint8)
        lengths = np.ascontiguousarray(lengths, dtype=np.int64)
        self.n, self.lmax = codes.shape
        self.nw = self.lmax // 32 + 4
        self.lengths = lengths
        self.device = device
        # codes and int32 lengths staged in one pinned host buffer and sent with
        # one asynchronous copy on the stream the packing runs on (the pinned
        # buffer stays referenced with this object)
        nc = -(-codes.size // 4) * 4
        self._host = torch.empty(nc + 4 * self.n, dtype=torch.uint8, pin_memory=True)
        hb = self._host.numpy()
        hb[: codes.size] = codes.reshape(-1)
        hb[nc:].view(np.int32)[:] = lengths
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            buf = self._host.to(device, non_blocking=True)
        self.codes = buf[: codes.size].view(self.n, self.lmax)
        self.len = buf[nc:].view(torch.int32)
        self._pack(rc, stream)

    @classmethod
    def from_strings(cls, strings, device, *, rc: bool = True, stream=None, min_width: int = 1,
                     staging: torch.Tensor | None = None):
        """(DeviceSeqs of DNA strings, bad [n] int32 on the device): the
        strings' bytes, unpadded, go over in one pinned copy and
        gkm_decode_ascii writes the padded codes (width max(Lmax, min_width); N
        as A) and flags rows with a letter other than ACGTN (bad = 1; empty
        rows are the caller's to flag, from .lengths).  staging: a pinned
        uint8 host buffer to stage the bytes in when it is large enough (the
        caller keeps it untouched until the copy on `stream` completes);
        otherwise a new pinned buffer is allocated."""
        self = cls.__new__(cls)
        n = len(strings)
        lens = np.fromiter(map(len, strings), dtype=np.int64, count=n)
        # one byte per character: a non-ASCII character becomes "?", flagged
        raw = "".join(strings).encode("ascii", "replace")
        self.n, self.lmax = n, max(int(lens.max()) if n else 0, min_width, 1)
        self.nw = self.lmax // 32 + 4
        self.lengths = lens
        self.device = device
        nr = -(-len(raw) // 8) * 8
        if staging is not None and staging.numel() >= nr + 12 * n:
            self._host = staging[: nr + 12 * n]
        else:
            self._host = torch.empty(nr + 12 * n, dtype=torch.uint8, pin_memory=True)
        hb = self._host.numpy()
        hb[: len(raw)] = np.frombuffer(raw, dtype=np.uint8)
        starts = hb[nr: nr + 8 * n].view(np.int64)
        if n:
            starts[0] = 0
            np.cumsum(lens[:-1], out=starts[1:])
        hb[nr + 8 * n:].view(np.int32)[:] = lens
        with torch.cuda.stream(stream) if stream is not None else contextlib.nullcontext():
            buf = self._host.to(device, non_blocking=True)
        self.len = buf[nr + 8 * n:].view(torch.int32)
        self.codes = torch.empty((n, self.lmax), dtype=torch.uint8, device=device)
        bad = torch.empty(n, dtype=torch.int32, device=device)
        _lib.call("gkm_decode_ascii", _p(buf), _p(buf[nr:]), _p(self.len), n, self.lmax,
                  _p(self.codes), _p(bad), stream_handle(stream))
        self._pack(rc, stream)
        return self, bad

    def _pack(self, rc: bool, stream):
        self.planes = torch.empty((self.n, self.nw, 2), dtype=torch.int32, device=self.device)
        self.planes_rc = torch.empty_like(self.planes) if rc else None
        _lib.call("gkm_pack_planes", _p(self.codes), _p(self.len), self.n, self.lmax, self.nw,
                  _p(self.planes), _p(self.planes_rc), stream_handle(stream))
        self._windows = {}

    def windows(self, l: int, stream=None):
        """(fwd, rc) packed l-mer windows [n, Lmax - l + 1] int32 (uint32 bits)."""
        if l not in self._windows:
            W = self.lmax - l + 1
            if W < 1:  # no row holds an l-mer: one column of PAD (0xFFFFFFFF)
                fwd = torch.full((self.n, 1), -1, dtype=torch.int32, device=self.device)
                self._windows[l] = (fwd, fwd.clone())
                return self._windows[l]
            fwd = torch.empty((self.n, W), dtype=torch.int32, device=self.device)
            rcw = torch.empty_like(fwd)
            _lib.call("gkm_pack_windows", _p(self.codes), _p(self.len), self.n, self.lmax, l,
                      _p(fwd), _p(rcw), stream_handle(stream))
            self._windows[l] = (fwd, rcw)
        return self._windows[l]

    @property
    def max_windows(self) -> int:
        return int(self.lengths.max()) if self.n else 0


class KernelParams:
    """What the kernels need of a GkmKernel, on one device."""

    def __init__(self, kernel: GkmKernel, device):
        base = kernel.base
        if not isinstance(base, DirectGkmKernel):
            raise TypeError(f"unsupported kernel {type(base).__name__}")
        self.kernel = kernel
        self.base = base
        self.l = base.l
        self.dmax = base.dmax
        self.nm = max(self.dmax + 1, 1)
        self.rc = base.include_rc
        self.normalize = base.normalize
        self.rbf = kernel.gamma is not None
        self.gamma = float(kernel.gamma) if self.rbf else 0.0
        self.device = device
        self.h = base.h_padded()
        self.tab = torch.tensor(self.h, dtype=torch.float64, device=device)
        self.weighted = base.weight_table(1) is not None
        self._wtab = None
        self._wtab_n = 0
        self._wtab_old = []

    def wtab(self, max_len: int):
        """(uint8 weight-by-distance table on the device covering sequences up to
        max_len bases, its length) or (None, 0)."""
        if not self.weighted:
            return None, 0
        need = max(1, max_len - self.l + 1) // 2 + 1
        if self._wtab is None or self._wtab_n < need:
            n = max(need, 1024)
            t = self.base.weight_table(2 * (n - 1))
            if self._wtab is not None:
                # a kernel queued on another stream may still read the old
                # table: it is kept (a few KB) rather than freed
                self._wtab_old.append(self._wtab)
            self._wtab = torch.from_numpy(np.ascontiguousarray(t)).to(self.device)
            self._wtab_n = len(t)
        return self._wtab, self._wtab_n


def group_bounds(lengths: np.ndarray) -> np.ndarray:
    """[ceil(n/64), 2] int32 (min, max) length of each 64-sequence group."""
    n = len(lengths)
    g = (n + WAVE - 1) // WAVE
    pad = np.full(g * WAVE, -1, dtype=np.int64)
    pad[:n] = lengths
    blk = pad.reshape(g, WAVE)
    mx = blk.max(axis=1)
    mn = np.where(blk < 0, np.iinfo(np.int64).max, blk).min(axis=1)
    return np.stack([mn, mx], axis=1).astype(np.int32)


def launch_shape(B: int, S: int, norient: int, n_cu: int = 256):
    """(rows_per_block, nsplit) for gkm_cross_hist: enough waves to fill the
    chip (>= ~16 per CU), each block's 4 waves busy, and grid.y <= 65535."""
    groups = max(1, (S + WAVE - 1) // WAVE)
    target_units = 16 * 4 * n_cu  # waves of work
    per_row = groups * norient
    nsplit = int(min(16, max(1, math.ceil(target_units / max(1, B * per_row)))))
    units_per_row = norient * nsplit
    rows_per_block = max(1, min(B, math.ceil(4 / units_per_row)))
    rows_per_block = max(rows_per_block, math.ceil(B / 65535))
    return rows_per_block, nsplit


class KernelTimer:
    """HIP-event timings of the hot kernels' launches (events recorded on the
    stream each launch goes to) with each launch's algorithmic work: window-pair
    comparisons (k_cross_hist, k_self_hist, k_sweep) or bytes (k_index_tile,
    k_table_score, k_table_explain)."""

    def __init__(self):
        self.records = []

    def add(self, kind: str, ev0, ev1, units: float):
        self.records.append((kind, ev0, ev1, units))

    def summary(self) -> dict:
        """kind -> {launches, total_ms, mean_ms, units_per_launch}."""
        torch.cuda.synchronize()
        out = {}
        for kind, a, b, u in self.records:
            d = out.setdefault(kind, {"launches": 0, "total_ms": 0.0, "units": 0.0})
            d["launches"] += 1
            d["total_ms"] += a.elapsed_time(b)
            d["units"] += u
        for d in out.values():
            d["mean_ms"] = d["total_ms"] / d["launches"]
            d["units_per_launch"] = d.pop("units") / d["launches"]
        return out


TIMER: KernelTimer | None = None


def _timed(kind: str, units_fn, stream, launch):
    """Run launch(); with TIMER set, bracket it with events on the stream."""
    timer = TIMER
    if timer is None:
        return launch()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    r = launch()
    ev1 = torch.cuda.Event(enable_timing=True)
    ev1.record(stream)
    timer.add(kind, ev0, ev1, units_fn())
    return r


def comparisons(kp: KernelParams, x_lengths, y_lengths, norient: int) -> int:
    """Window-pair comparisons of rows x against columns y: sum_b sum_s
    norient * Wx_b * Wy_s (W = windows of a sequence)."""
    wx = np.clip(np.asarray(x_lengths, dtype=np.int64) - kp.l + 1, 0, None)
    wy = np.clip(np.asarray(y_lengths, dtype=np.int64) - kp.l + 1, 0, None)
    return int(norient * wx.sum() * wy.sum())


def cross_hist(kp: KernelParams, xs: DeviceSeqs, x0: int, x1: int, ys: DeviceSeqs,
               ygroups: torch.Tensor, *, col_min: int = -1, out=None, shape=None,
               stream=None) -> torch.Tensor:
    """[x1 - x0, S, nm] int64 histograms of rows x0..x1 of xs against ys."""
    B = x1 - x0
    S = ys.n
    if out is None:
        out = torch.empty((B, S, kp.nm), dtype=torch.int64, device=kp.device)
    _lib.call("gkm_zero_u64", _p(out), out.numel(), stream_handle(stream))
    norient = 2 if (kp.rc and ys.planes_rc is not None) else 1
    rpb, nsplit = shape if shape is not None else launch_shape(B, S, norient)
    wt, wn = kp.wtab(max(xs.lmax, ys.lmax))
    xpl = xs.planes[x0:x1]
    _timed("k_cross_hist", lambda: comparisons(kp, xs.lengths[x0:x1], ys.lengths, norient),
           stream, lambda: _lib.call(
               "gkm_cross_hist", _p(xpl), _p(xs.len[x0:x1]), B, xs.nw, _p(ys.planes),
               _p(ys.planes_rc) if norient == 2 else None, _p(ys.len), _p(ygroups), S, ys.nw,
               kp.l, kp.nm, _p(wt), wn, rpb, nsplit, col_min, x0, _p(out),
               stream_handle(stream)))
    return out


def self_kraw(kp: KernelParams, xs: DeviceSeqs, stream=None) -> torch.Tensor:
    """K_raw(x, x) [n] float64 of every sequence (forward x forward and RC)."""
    hist = torch.empty((xs.n, kp.nm), dtype=torch.int64, device=kp.device)
    _lib.call("gkm_zero_u64", _p(hist), hist.numel(), stream_handle(stream))
    wt, wn = kp.wtab(xs.lmax)
    longest = xs.max_windows
    # enough diagonal chunks that a lone long sequence still spreads over CUs
    nd = 2 * (2 * max(longest - kp.l, 0) + 1)
    nchunks = max(1, min(64, math.ceil(nd / 256), math.ceil(4096 / max(1, xs.n))))
    norient = 2 if kp.rc else 1
    _timed("k_self_hist",
           lambda: float(norient * (np.clip(xs.lengths - kp.l + 1, 0, None) ** 2).sum()), stream,
           lambda: _lib.call(
               "gkm_self_hist", _p(xs.planes), _p(xs.planes_rc) if kp.rc else None, _p(xs.len),
               xs.n, xs.nw, kp.l, kp.nm, _p(wt), wn, nchunks, _p(hist), stream_handle(stream)))
    out = torch.empty(xs.n, dtype=torch.float64, device=kp.device)
    _lib.call("gkm_combine_kraw", _p(hist), xs.n, kp.nm, _p(kp.tab), _p(out),
              stream_handle(stream))
    return out


def neighbor_masks(l: int, table, dmax: int):
    """(masks uint32, levels uint8) of every l-mer XOR pattern with m <= dmax
    non-zero 2-bit digits whose pair weight table[m] is non-zero, sorted by
    mask (kernels/lmer_index.py:67-87 enumerates the same set)."""
    from itertools import combinations, product

    out = []
    for m in range(dmax + 1):
        if table[m] == 0:
            continue
        for pos in combinations(range(l), m):
            for digits in product((1, 2, 3), repeat=m):
                v = 0
                for p_, d_ in zip(pos, digits):
                    v |= d_ << (2 * p_)
                out.append((v, m))
    out.sort()
    masks = np.array([v for v, _ in out], dtype=np.uint32)
    levels = np.array([m for _, m in out], dtype=np.uint8)
    return masks, levels


def paired_masks(masks: np.ndarray, levels: np.ndarray, tables: int):
    """The index query's order of the neighbourhood masks: pairs of masks that
    differ in bit 0 of one digit q < tables, each pair at lanes 2k, 2k + 1 of a
    wave (so both probes read one 128-byte line of record table q, which swaps
    the l-mer's digits 0 and q), then the unpaired masks; levels[t] | q << 4.
    A greedy matching, masks with the fewest free partners first (l = 11, d = 3:
    2,432 pairs of the 4,984 masks)."""
    n = len(masks)
    where = {int(m): i for i, m in enumerate(masks)}
    adj = [[(where[int(m) ^ (1 << (2 * q))], q) for q in range(tables)
            if int(m) ^ (1 << (2 * q)) in where] for m in masks]
    mate = [-1] * n
    table = [0] * n
    for i in sorted(range(n), key=lambda i: (len(adj[i]), int(masks[i]))):
        if mate[i] >= 0:
            continue
        best = None
        for j, q in adj[i]:
            if mate[j] < 0:
                free = sum(1 for k, _ in adj[j] if mate[k] < 0)
                if best is None or (free, q) < best[0]:
                    best = ((free, q), j, q)
        if best is not None:
            j, q = best[1], best[2]
            mate[i], mate[j] = j, i
            table[i] = table[j] = q
    pairs = sorted({(min(i, mate[i]), max(i, mate[i])) for i in range(n) if mate[i] >= 0},
                   key=lambda pr: (table[pr[0]], int(masks[pr[0]])))
    sel = np.array([k for pr in pairs for k in pr] + [i for i in range(n) if mate[i] < 0],
                   dtype=np.int64)
    lv = (levels.astype(np.uint8) | (np.array(table, dtype=np.uint8) << 4))[sel]
    return masks[sel], lv


# Throughput of the two scoring paths on one MI355X (measured, l=11 d=3, 300-1000
# bp rows, 500-72,000 SVs), for choosing between them: window-pair comparisons
# per second of gkm_cross_hist; (window, mask, SV tile) probes and index
# entries per second of gkm_index_hist.
BRUTE_COMPARISONS_PER_S = 1.8e13
INDEX_PROBES_PER_S = 1.0e11          # unweighted: bucket records
INDEX_PROBES_PER_S_WEIGHTED = 5.0e10  # weighted: bounds, then ids and weights
INDEX_ENTRIES_PER_S = 2.0e12
INDEX_MAX_L = 12


class DeviceIndex:
    """The l-mers of every SV window (forward and RC) as a CSR in HBM, split by
    SV tile: keys lmer * n_tiles + tile, offsets [4^l n_tiles + 1] uint32, ids
    uint16 (SV within its tile), vals uint8 window weights (-t 4/-t 5); the
    neighbourhood masks and their levels.  kernels/lmer_index.py:161-200 builds
    the same index without the tiles (a tile's counts fit one block's LDS)."""

    def __init__(self, kp: KernelParams, seqs: DeviceSeqs, stream=None, tile: int | None = None,
                 count_bits: int | None = None):
        l = kp.l
        nl = 1 << (2 * l)
        dev = kp.device
        S = seqs.n
        # packed LDS counts (unweighted, and a chunk of at least one window
        # keeps every count below 2^16): 16-bit, or 8-bit with the kernel's
        # exact 16-bit fallback where nm >= 3 makes them take fewer words;
        # count_bits (32, 16, 8) forces a width where it applies
        nb = 2 if kp.rc else 1
        sv_windows = int(np.clip(seqs.lengths - l + 1, 0, None).max()) * nb if S else 0
        self.pack16 = not kp.weighted and 0 < sv_windows < 65536 and count_bits != 32
        if kp.weighted:  # 16-bit weighted counts (carry-checked, 32-bit fallback): uint32 partials
            self.pack = 1 if count_bits == 32 else 2
        elif not self.pack16:
            self.pack = 1
        else:
            self.pack = 2 if count_bits == 16 else index_pack(kp.nm, True)
        self.tile = tile or min(index_tile(kp.nm, self.pack, S), max(256, -(-S // 256) * 256))
        self.n_tiles = (S + self.tile - 1) // self.tile
        fwd, rcw = seqs.windows(l, stream)
        blocks = [fwd] + ([rcw] if kp.rc else [])
        keys = torch.cat(blocks, dim=1).to(torch.int64) & 0xFFFFFFFF  # [S, nb*W]
        nw = keys.shape[1]
        W = fwd.shape[1]
        flat = keys.reshape(-1)
        valid = flat < nl
        sv = torch.arange(S, device=dev, dtype=torch.int64).repeat_interleave(nw)
        vals = None
        if kp.weighted:
            wt, _ = kp.wtab(seqs.lmax)
            ws = torch.from_numpy(np.clip(seqs.lengths - l + 1, 0, None)).to(dev)[:, None]
            j = torch.arange(W, device=dev)[None, :]
            d = [(ws // 2 - j).abs()] + ([(ws // 2 - (ws - 1 - j)).abs()] if kp.rc else [])
            dist = torch.cat(d, dim=1).reshape(-1).clamp(0, wt.numel() - 1)
            vals = wt[dist][valid]
        flat = flat[valid]
        sv = sv[valid]
        key = flat * self.n_tiles + sv // self.tile
        order = torch.argsort(key, stable=True)
        self.ids = (sv[order] % self.tile).to(torch.int16).contiguous()  # uint16 bits
        self.vals = None if vals is None else vals[order].contiguous()
        counts = torch.bincount(key, minlength=nl * self.n_tiles)
        offsets = torch.zeros(nl * self.n_tiles + 1, dtype=torch.int64, device=dev)
        offsets[1:] = torch.cumsum(counts, 0)
        self.n_entries = int(offsets[-1].item())
        if self.n_entries >= 2**31:
            raise MemoryError("l-mer index over 2^31 entries")
        self.offsets = offsets.to(torch.int32)  # uint32 bits
        # bucket records, 32 bytes per key (the l-mer index query) -- the count
        # and ids in one load, the n_tiles records of an l-mer in one cache
        # line.  Unweighted: halfword 0 = min(count, 0xFFFF); a bucket of at
        # most RECORD_IDS entries holds its ids in halfwords 1..count, a longer
        # one its first RECORD_IDS_BIG ids in halfwords 1..11, its start in
        # word 6 and its count in word 7.  Weighted: start, count, the first
        # RECORD_IDS_W ids (halfwords 4..11) and their uint8 weights (bytes
        # 24..31).
        self.records = None
        nk = nl * self.n_tiles
        ks = key[order]
        if (self.pack16 or self.vals is not None) and nk * 32 <= RECORD_MAX_BYTES:
            rec = torch.zeros((nk, 16), dtype=torch.int16, device=dev)
            r32 = rec.view(torch.int32)
            cnt = offsets[1:] - offsets[:-1]
            rank = torch.arange(self.n_entries, device=dev) - offsets[ks]
            if self.vals is None:
                big = cnt > RECORD_IDS
                h0 = cnt.clamp(max=0xFFFF)
                rec[:, 0] = (h0 - (h0 >= 0x8000).to(torch.int64) * 0x10000).to(torch.int16)
                rec[:, 1:] = -1  # 0xFFFF past the ids: the sink mark
                cap = torch.where(big, RECORD_IDS_BIG, RECORD_IDS)
                sel = rank < cap[ks]
                rec.view(-1)[ks[sel] * 16 + 1 + rank[sel]] = self.ids[sel]
                r32[big, 6] = self.offsets[:-1][big]
                r32[big, 7] = cnt[big].to(torch.int32)
                del big, h0, cap
            else:
                r32[:, 0] = self.offsets[:-1]
                r32[:, 1] = cnt.to(torch.int32)
                sel = rank < RECORD_IDS_W
                rec.view(-1)[ks[sel] * 16 + 4 + rank[sel]] = self.ids[sel]
                rb = rec.view(torch.uint8)
                rb.view(-1)[ks[sel] * 32 + 24 + rank[sel]] = self.vals[sel].to(torch.uint8)
            self.records = r32
            del rec, rank, sel, cnt
        # l-mer records, 64 bytes per l-mer (GkmExplain through the index, which
        # sums over every tile): unweighted with at most LREC_MAX_TILES tiles,
        # whose buckets of one l-mer are consecutive -- the run's start, the
        # tiles' counts as bytes (255: saturated), the run's first LREC_IDS ids
        self.lmer_records = None
        if self.pack16 and self.n_tiles <= LREC_MAX_TILES and nl * 64 <= RECORD_MAX_BYTES:
            rec = torch.zeros((nl, 32), dtype=torch.int16, device=dev)
            r32 = rec.view(torch.int32)
            r32[:, 0] = self.offsets[:-1:self.n_tiles][:nl]
            cnt = (offsets[1:] - offsets[:-1]).clamp(max=255).to(torch.uint8).view(nl, self.n_tiles)
            rec.view(torch.uint8)[:, 4:4 + self.n_tiles] = cnt
            lmer = ks // self.n_tiles
            rank = torch.arange(self.n_entries, device=dev) - offsets[lmer * self.n_tiles]
            sel = rank < LREC_IDS
            rec.view(-1)[lmer[sel] * 32 + 4 + rank[sel]] = self.ids[sel]
            self.lmer_records = r32
            del rec, lmer, rank, sel, cnt
        del ks
        # INDEX_PAD entries past the last bucket: the query reads ahead of a
        # bucket's overflow (and reads the first entries where a lane has none)
        self.ids = torch.cat([self.ids, torch.zeros(INDEX_PAD, dtype=torch.int16, device=dev)])
        if self.vals is not None:
            self.vals = torch.cat([self.vals,
                                   torch.zeros(INDEX_PAD, dtype=self.vals.dtype, device=dev)])
        # record tables 1..q: the bucket records again with the l-mer's digits 0
        # and q swapped, so a pair of masks differing in bit 0 of digit q reads
        # one 128-byte line (with at most 2 tiles an l-mer pair's records of a
        # tile share one); tables beyond RECORD_TABLES_MAX_BYTES are not built
        self.record_tables = 1
        if self.records is not None and self.n_tiles <= 2:
            tables = int(min(l, max(1, RECORD_TABLES_MAX_BYTES // (nk * 32))))
            if tables > 1:
                base = self.records.view(nl, self.n_tiles * 8)
                allr = torch.empty((tables, nl, self.n_tiles * 8), dtype=torch.int32, device=dev)
                allr[0] = base
                lm = torch.arange(nl, device=dev, dtype=torch.int64)
                for q in range(1, tables):
                    x = (lm ^ (lm >> (2 * q))) & 3
                    allr[q] = base[lm ^ x ^ (x << (2 * q))]
                del lm, x, base
                self.records = allr.view(tables * nk, 8)  # table 0 first: key order
                self.record_tables = tables
        masks, levels = neighbor_masks(l, kp.h, kp.dmax)
        if self.record_tables > 1:
            masks, levels = paired_masks(masks, levels, self.record_tables)
        self.masks = torch.from_numpy(masks.view(np.int32)).to(dev)
        self.levels = torch.from_numpy(levels).to(dev)
        self.n_masks = len(masks)
        self.per_lmer = self.n_entries / nl
        self.max_weight = int(kp.base.weight_table(1)[0]) if kp.weighted else 1
        self.max_sv_windows = int(np.clip(seqs.lengths - l + 1, 0, None).max()) * len(blocks)
        del keys, flat, sv, order, counts, key

    def nbytes(self) -> int:
        t = [self.ids, self.offsets, self.masks, self.levels]
        if self.vals is not None:
            t.append(self.vals)
        t += [r for r in (self.records, self.lmer_records) if r is not None]
        return sum(x.numel() * x.element_size() for x in t)


def index_pack(nm: int, unweighted: bool) -> int:
    """Counts per 32-bit LDS word of the index query: 2 (16-bit; weighted
    counts with the kernel's carry check and 32-bit fallback) or 4 (8-bit,
    unweighted, where that takes fewer words than 16-bit: nm >= 3)."""
    if not unweighted:
        return 2
    return 4 if (nm + 3) // 4 < (nm + 1) // 2 else 2


# LDS words a tile's counts may take: the kernels' 159 KiB for counts and the
# records' 64 sink words (the rest of the CU's 160 KiB is their static LDS)
INDEX_LDS_WORDS = 159 * 1024 // 4 - 64


def index_tile(nm: int, pack: int | bool = 1, S: int | None = None) -> int:
    """SVs per index tile: the tile's counts ([tile, nm] uint32, or `pack`
    counts per word; True: the unweighted packing index_pack picks) take 144
    KiB of the CU's 160 KiB of LDS -- the index is probe-bound, so the
    largest tile that fits wins (fewer tiles, fewer probes); at most 65,536
    (uint16 ids).  Given the SV count S, a tile up to INDEX_LDS_WORDS is taken
    where that needs fewer tiles (20,000 SVs with weighted 16-bit counts: one
    tile of 20,224 instead of two of 18,432), the SVs spread evenly."""
    if pack is True or pack is False:
        pack = index_pack(nm, pack)
    words = -(-nm // pack)
    tile = min(65536, max(256, (36864 // words) // 256 * 256))
    if S is None or S <= tile:
        return tile
    big = min(65536, (INDEX_LDS_WORDS // words) // 256 * 256)
    n = -(-S // big)
    if n < -(-S // tile):
        return -(-(-(-S // n)) // 256) * 256
    return tile


MAX_PARTIAL_BYTES = 8 << 30  # per-chunk index counts of one launch (rows per launch)
LREC_IDS = 28                # ids held in a 64-byte l-mer record
LREC_MAX_TILES = 4           # tiles an l-mer record covers
RECORD_IDS = 15              # ids an unweighted bucket record holds (buckets up to 15)
RECORD_IDS_BIG = 11          # ids it holds beside the start and count (longer buckets)
INDEX_PAD = 16               # entries past the last bucket of ids and vals (GKM_INDEX_PAD)
RECORD_IDS_W = 8             # (id, weight) pairs held in a record of a weighted index
RECORD_MAX_BYTES = 4 << 30   # larger record tables (l >= 13 with many tiles) are not built
RECORD_TABLES_MAX_BYTES = 4 << 30  # all the digit-swapped record tables together


def index_hist(kp: KernelParams, idx: DeviceIndex, xs: DeviceSeqs, x0: int, x1: int, S: int, *,
               col_min: int = -1, out=None, stream=None,
               target_blocks: int | None = None) -> torch.Tensor:
    """[x1 - x0, S, nm] int64 histograms of rows x0..x1 through the index (in
    row groups whose per-chunk partial counts stay under MAX_PARTIAL_BYTES)."""
    B = x1 - x0
    if out is None:
        out = torch.empty((B, S, kp.nm), dtype=torch.int64, device=kp.device)
    _lib.call("gkm_zero_u64", _p(out), out.numel(), stream_handle(stream))
    fwd, _ = xs.windows(kp.l, stream)
    Wx = fwd.shape[1]
    # rows per launch: the per-chunk partials of the smallest chunk (16
    # windows, or the 16-bit cap) must fit the budget
    min_cw = 16 if not idx.pack16 else max(1, min(16, 65535 // max(1, idx.max_sv_windows)))
    per_row = -(-Wx // min_cw) * S * kp.nm * (2 if idx.pack16 else 4)
    step = max(1, min(B, MAX_PARTIAL_BYTES // max(1, per_row), 65535))
    queue = col_min < 0 and target_blocks is None
    for r0 in range(0, B, step):
        r1 = min(B, r0 + step)
        if queue:  # one persistent launch per row group (the Gram's triangle: a grid)
            _index_launch_queue(kp, idx, xs, x0 + r0, x0 + r1, S, out[r0:r1], stream)
        else:
            _index_launch(kp, idx, xs, x0 + r0, x0 + r1, S, col_min, out[r0:r1], stream,
                          target_blocks)
    return out


INDEX_CHUNK_UNIFORM = 32  # query windows per block, rows of one length
INDEX_CHUNK_RAGGED = 16   # rows of very different lengths
INDEX_CHUNK_MIN = 16      # the finest chunk: one window per wave of a block
INDEX_MIN_BLOCKS = 128    # finer chunks below this many blocks (256 CUs)


def index_blocks(x_lengths, l: int, n_tiles: int = 1, chunk: int | None = None) -> int:
    """Block target of an index launch that gives chunks of about `chunk`
    query windows -- by default (measured on 1-105 row batches of 50-2000 bp
    against 20k and 72k SVs) 32 for rows of one length, 16 when the launch
    would have fewer than INDEX_MIN_BLOCKS blocks (a single 1 kb query: 0.57
    -> 0.35 ms), and 16 for rows of very different lengths, which balances
    the CUs.  Finer chunks would leave waves of a block idle."""
    wx = np.clip(np.asarray(x_lengths, dtype=np.int64) - l + 1, 0, None)
    per = max(1, n_tiles) * len(wx)
    w = max(1, int(wx.max()))
    if chunk is None:
        if w > 1.5 * max(1.0, wx.mean()):
            chunk = INDEX_CHUNK_RAGGED
        else:
            chunk = INDEX_CHUNK_UNIFORM
            while chunk > INDEX_CHUNK_MIN and per * -(-w // chunk) < INDEX_MIN_BLOCKS:
                chunk //= 2
    return per * max(1, -(-w // chunk))


class IndexLaunch:
    """The arguments of one gkm_index_hist call over rows x0..x1 (window
    chunking, sorted query windows, partial-count format); `args` are its
    C-ABI arguments up to (not including) partial."""

    def __init__(self, kp, idx, xs, x0, x1, S, col_min, stream, target_blocks, chunk=None):
        B = x1 - x0
        if target_blocks is None and chunk is None:
            target_blocks = index_blocks(xs.lengths[x0:x1], kp.l, idx.n_tiles)
        fwd, _ = xs.windows(kp.l, stream)
        Wx = fwd.shape[1]
        wx_max = max(1, int(np.clip(xs.lengths[x0:x1] - kp.l + 1, 0, None).max()))
        if chunk is not None:  # windows per block given (the persistent form)
            cw = max(1, int(chunk))
        else:
            nchunks = max(1, min(-(-target_blocks // (idx.n_tiles * B)),
                                 -(-wx_max // INDEX_CHUNK_MIN)))
            cw = -(-wx_max // nchunks)
        if idx.vals is not None:  # per-block uint32 counts must not overflow
            cap = (2**32 - 1) // (idx.max_weight**2 * max(1, idx.max_sv_windows))
            cw = max(1, min(cw, cap))
        elif idx.pack16:  # nor the 16-bit ones: a window meets an SV window at most once
            cw = max(1, min(cw, 65535 // max(1, idx.max_sv_windows)))
        self.B, self.Wx, self.cw = B, Wx, cw
        self.nchunks = -(-Wx // cw)
        self.pack_code = {1: 0, 2: 1, 4: 2}[idx.pack]
        # bytes, uint16 or uint32 partials (gkm_index_reduce's count_bytes)
        self.count_bytes = 1 if idx.pack == 4 else 2 if idx.pack16 else 4
        wt, _ = kp.wtab(xs.lmax)
        # each row's windows in l-mer order (PAD last): a block's windows then
        # share high digits, and so do their neighbours' offsets -- cache lines
        # reused (32-bit keys: PAD, 0xFFFFFFFF, sorts last as INT32_MAX, >= 4^l)
        keys = fwd[x0:x1]
        keys = torch.where(keys < 0, torch.full_like(keys, 2**31 - 1), keys)
        keys, pos = torch.sort(keys, dim=1)
        self.swin = keys.contiguous()
        self.spos = pos.to(torch.int32).contiguous() if wt is not None else None
        self.args = (_p(self.swin), _p(self.spos), _p(xs.len[x0:x1]), B, Wx, _p(idx.offsets),
                     _p(idx.ids), _p(idx.vals), _p(idx.records), idx.record_tables,
                     _p(idx.masks), _p(idx.levels), idx.n_masks, kp.l, kp.nm, _p(wt), S,
                     idx.tile, idx.n_tiles, cw, self.pack_code, col_min, x0)

    def partial_bytes(self, kp, idx, S) -> int:
        n_all = self.nchunks * self.B * S * kp.nm
        if self.count_bytes == 1:
            return 2 * n_all + self.nchunks * self.B * idx.n_tiles
        return n_all * self.count_bytes


def _index_launch(kp, idx, xs, x0, x1, S, col_min, out, stream, target_blocks):
    L = IndexLaunch(kp, idx, xs, x0, x1, S, col_min, stream, target_blocks)
    partial = torch.empty(L.partial_bytes(kp, idx, S), device=kp.device, dtype=torch.uint8)
    # the query (timed on its own) and the sum of its per-chunk partials
    _timed("k_index_tile", lambda: index_bytes(kp, idx, xs.lengths[x0:x1]), stream,
           lambda: _lib.call("gkm_index_hist", *L.args, _p(partial), None,
                             stream_handle(stream)))
    _lib.call("gkm_index_reduce", _p(partial), L.count_bytes, _p(xs.len[x0:x1]),
              L.B, L.Wx, kp.l, kp.nm, S, idx.tile, L.cw, _p(out), stream_handle(stream))
    return out


def _index_launch_queue(kp, idx, xs, x0, x1, S, out, stream):
    """Rows x0..x1's histograms added into out through one persistent launch
    (gkm_index_hist_queue): chunks of INDEX_CHUNK_UNIFORM windows for rows of
    one length, INDEX_CHUNK_RAGGED otherwise."""
    wx = np.clip(xs.lengths[x0:x1] - kp.l + 1, 0, None)
    ragged = int(wx.max(initial=0)) > 1.5 * max(1.0, float(wx.mean()) if len(wx) else 1.0)
    chunk = INDEX_CHUNK_RAGGED if ragged else INDEX_CHUNK_UNIFORM
    L = IndexLaunch(kp, idx, xs, x0, x1, S, -1, stream, None, chunk=chunk)
    n_pairs = int(np.minimum(-(-wx // L.cw), L.nchunks).sum())
    ws = torch.empty(_lib.load().gkm_index_queue_bytes(L.B, 0, idx.n_tiles, kp.nm),
                     dtype=torch.uint8, device=kp.device)
    partial = torch.empty(L.partial_bytes(kp, idx, S), device=kp.device, dtype=torch.uint8)
    _timed("k_index_tile_queue", lambda: index_bytes(kp, idx, xs.lengths[x0:x1]), stream,
           lambda: _lib.call("gkm_index_hist_queue", *L.args[:-2], n_pairs, n_cus(kp.device),
                             _p(ws), _p(partial), _p(out), stream_handle(stream)))
    return out


INDEX_PROBE_BYTES = 16.0  # algorithmic bytes per (window, mask) probe: bucket offsets
INDEX_HIT_BYTES = 12.0    # per hit: a 4-byte id and an 8-byte FP64 read-modify-write


def index_bytes(kp: KernelParams, idx: DeviceIndex, x_lengths) -> float:
    """Algorithmic bytes of an index query (the reference algorithm's traffic,
    kernels/lmer_index_gpu.py:9-39): 16 per (window, mask) probe, 12 per hit
    (SV window equal to the probed l-mer)."""
    wx = float(np.clip(np.asarray(x_lengths) - kp.l + 1, 0, None).sum())
    probes = wx * idx.n_masks
    return probes * INDEX_PROBE_BYTES + probes * idx.per_lmer * INDEX_HIT_BYTES


class SVModel:
    """An SVM's support vectors resident on the GPU: bit-planes sorted by
    length, 64-SV group bounds, coefficients and self-kernels in that order."""

    def __init__(self, kp: KernelParams, codes: np.ndarray, lengths: np.ndarray,
                 coef: np.ndarray, stream=None):
        self.kp = kp
        order = np.argsort(lengths, kind="stable")
        self.order = order
        self.seqs = DeviceSeqs(codes[order], lengths[order], kp.device, rc=kp.rc, stream=stream)
        self.groups = torch.from_numpy(group_bounds(lengths[order])).to(kp.device)
        self.coef = torch.from_numpy(np.ascontiguousarray(coef[order], dtype=np.float64)).to(
            kp.device)
        self.S = len(order)
        if kp.normalize:
            self.ds = self_kraw(kp, self.seqs, stream)
        else:
            self.ds = torch.ones(self.S, dtype=torch.float64, device=kp.device)
        # exact scoring of kernels without RBF: fixed-point SV weights
        self.wfix, self.fix_shift = (None, 0) if kp.rbf else fixed_weights(
            coef[order], self.ds.cpu().numpy() if kp.normalize else None, kp.device)

        self._index = None
        # exact l-mer tables of non-RBF models (tables.py), built on demand:
        # "V", "score", ("explain", perturbation)
        self.tables: dict = {}

    def index(self):
        """The l-mer index of the SVs (built on first use), or None when l is
        too large for 4^l offsets."""
        if self._index is None and self.kp.l <= INDEX_MAX_L and self.kp.dmax >= 0:
            self._index = DeviceIndex(self.kp, self.seqs)
        return self._index

    def _path_seconds(self, n_windows: float):
        """(index, brute-force) estimated GPU seconds for n_windows query
        windows against every SV (index None when it does not apply)."""
        norient = 2 if self.kp.rc else 1
        ws = float(np.clip(self.seqs.lengths - self.kp.l + 1, 0, None).sum())
        brute_s = n_windows * norient * ws / BRUTE_COMPARISONS_PER_S
        if self.kp.l > INDEX_MAX_L or self.kp.dmax < 0:
            return None, brute_s
        from math import comb

        n_masks = sum(comb(self.kp.l, m) * 3**m for m in range(self.kp.dmax + 1)
                      if self.kp.h[m] != 0)
        per_lmer = norient * ws / 4**self.kp.l
        probes = n_windows * n_masks
        n_tiles = -(-self.S // index_tile(self.kp.nm, not self.kp.weighted, self.S))
        rate = INDEX_PROBES_PER_S_WEIGHTED if self.kp.weighted else INDEX_PROBES_PER_S
        index_s = probes * n_tiles / rate + probes * per_lmer / INDEX_ENTRIES_PER_S
        return index_s, brute_s

    def query_seconds(self, n_windows: float) -> float:
        """Estimated GPU seconds of the cheaper per-SV path for n_windows query
        windows."""
        index_s, brute_s = self._path_seconds(float(n_windows))
        return brute_s if index_s is None else min(index_s, brute_s)

    def use_index_for(self, x_lengths, mode="auto") -> bool:
        """Whether scoring rows of these lengths goes through the index: forced
        by mode True/False, else the cheaper of the two paths by the cost model."""
        if mode is False or self.kp.l > INDEX_MAX_L or self.kp.dmax < 0:
            return False
        if mode is True:
            return True
        wx = np.clip(np.asarray(x_lengths) - self.kp.l + 1, 0, None).sum()
        index_s, brute_s = self._path_seconds(float(wx))
        return index_s < brute_s

    def diag_original_order(self) -> torch.Tensor:
        inv = np.empty_like(self.order)
        inv[self.order] = np.arange(len(self.order))
        return self.ds[torch.from_numpy(inv).to(self.ds.device)]

    def nbytes(self) -> int:
        t = [self.seqs.planes, self.seqs.codes, self.coef, self.ds, self.groups]
        if self.seqs.planes_rc is not None:
            t.append(self.seqs.planes_rc)
        return sum(x.numel() * x.element_size() for x in t)


FIX_BITS = 61  # the largest |fixed-point SV weight| is 2^61


def fixed_weights(coef: np.ndarray, ds: np.ndarray | None, device):
    """(wfix int64 [S] on the device, shift): the SV weights of exact scoring,
    w_s = coef_s / sqrt(K_raw(s, s)) (coef_s without normalisation; 0 for an
    SV without windows), as round(w_s 2^shift) with the largest at 2^61."""
    w = np.asarray(coef, dtype=np.float64).copy()
    if ds is not None:
        ds = np.asarray(ds, dtype=np.float64)
        w = np.where(ds > 0, w / np.sqrt(np.where(ds > 0, ds, 1.0)), 0.0)
    top = float(np.abs(w).max()) if w.size else 0.0
    shift = 0 if top == 0.0 else FIX_BITS - int(np.ceil(np.log2(top)))
    wfix = np.rint(np.ldexp(w, shift)).astype(np.int64)
    return torch.from_numpy(wfix).to(device), int(shift)


def index_score(sv: SVModel, xs: DeviceSeqs, r0: int, r1: int, dx, bias: float, out: torch.Tensor,
                stream=None, target_blocks: int | None = None) -> torch.Tensor:
    """out[r0:r1] = the rows' decision values straight from the l-mer index
    (gkm_index_score: each block reduces its counts to exact level sums; no
    histogram leaves the chip).  Kernels without RBF only."""
    kp, idx = sv.kp, sv.index()
    step = 65535
    for a in range(r0, r1, step):
        b = min(r1, a + step)
        L = IndexLaunch(kp, idx, xs, a, b, sv.S, -1, stream, target_blocks)
        psum = torch.empty(L.nchunks * L.B * idx.n_tiles * kp.nm * 16, dtype=torch.uint8,
                           device=kp.device)
        _timed("k_index_tile", lambda: index_bytes(kp, idx, xs.lengths[a:b]), stream,
               lambda: _lib.call(
                   "gkm_index_score", *L.args[:-2], _p(sv.wfix), _p(psum), _p(kp.tab),
                   _p(dx[a:b]) if dx is not None else None, sv.fix_shift, int(kp.normalize),
                   float(bias), _p(out[a:b]), stream_handle(stream)))
    return out


# query windows per work item of the persistent index launch: the serving
# batch's 103 ragged rows against 72,000 SVs took 6.26 / 6.12 / 5.89 / 5.90 ms
# at 16 / 20 / 24 / 32 (fewer items, each zeroing and folding a 147 KB LDS
# tile); 24 keeps the 8-bit counts further from their carry fallback than 32
QUEUE_CHUNK = 24


_N_CUS: dict = {}


def n_cus(device) -> int:
    """Compute units of the device (cached: the property query is not free)."""
    key = str(device)
    if key not in _N_CUS:
        _N_CUS[key] = int(torch.cuda.get_device_properties(device).multi_processor_count)
    return _N_CUS[key]


def index_score_queue(sv: SVModel, xs: DeviceSeqs, dx, bias: float, out: torch.Tensor,
                      stream=None, chunk: int = QUEUE_CHUNK, pairs: int = 0) -> torch.Tensor:
    """out[:] = every row's decision value from the l-mer index in one
    persistent launch (gkm_index_score_queue): the rows cut into work items
    of `chunk` windows (planned on the device from the uploaded lengths),
    popped by one block per CU from per-XCD queues.  The same bits as
    index_score.  Kernels without RBF.  pairs > 0: the last 4 * pairs rows
    are allele pairs in the layout of variant_rows, the alternate alleles
    (the last pairs rows) scored from the level sums of the other three
    blocks (gkm_index_score_queue_pairs) -- the same bits as probing them."""
    kp, idx = sv.kp, sv.index()
    if pairs < 0 or 4 * pairs > xs.n:
        raise ValueError(f"index_score_queue: {pairs} allele pairs need 4 rows each, "
                         f"{xs.n} rows")
    bq = xs.n - pairs
    L = IndexLaunch(kp, idx, xs, 0, bq, sv.S, -1, stream, None, chunk=chunk)
    # the chunk count the plan kernel computes from the lengths on the device
    n_pairs = int(np.minimum(-(-np.clip(xs.lengths[:bq] - kp.l + 1, 0, None) // L.cw),
                             L.nchunks).sum())
    ws = torch.empty(_lib.load().gkm_index_queue_bytes(bq, n_pairs, idx.n_tiles, kp.nm),
                     dtype=torch.uint8, device=kp.device)
    args = list(L.args[:-2])
    args[3] = xs.n  # every row's value; the plan and the probes take the first bq
    _timed("k_index_tile_queue", lambda: index_bytes(kp, idx, xs.lengths[:bq]), stream,
           lambda: _lib.call(
               "gkm_index_score_queue_pairs", *args, _p(sv.wfix), n_pairs, n_cus(kp.device),
               _p(ws), _p(kp.tab),
               _p(dx) if dx is not None else None, sv.fix_shift, int(kp.normalize), float(bias),
               _p(out), int(pairs), stream_handle(stream)))
    return out


# An allele pair is scored through its reference's level sums when the
# windows its variant changes are under this fraction of the allele's
VARIANT_ALIAS_FRACTION = 0.5


def variant_cut(ref: str, alt: str, l: int):
    """For one pair of allele strings: the base range (c0, c1) of the
    alternate's windows covering a base where the two differ (as
    variant_cuts), or None where the pair is better scored as two rows
    (lengths differ, identical, shorter than l, or changes spread over too
    many windows).  The differing bases come from the XOR of the strings'
    bytes read as one integer each: its highest set bit is in the first
    differing byte, its lowest in the last."""
    L = len(ref)
    if len(alt) != L or L < l:
        return None
    x = (int.from_bytes(ref.encode("ascii", "replace"), "big")
         ^ int.from_bytes(alt.encode("ascii", "replace"), "big"))
    if not x:
        return None
    lo = L - 1 - (x.bit_length() - 1) // 8
    hi = L - 1 - ((x & -x).bit_length() - 1) // 8
    W = L - l + 1
    w0, w1 = max(lo - l + 1, 0), min(hi, W - 1)
    if w1 - w0 + 1 >= VARIANT_ALIAS_FRACTION * W:
        return None
    return w0, w1 + l


def variant_cuts(lengths: np.ndarray, lo: np.ndarray, hi: np.ndarray, l: int):
    """For allele pairs of equal length (lengths) differing in bases lo..hi:
    (aliased mask, c0, c1) -- the pairs worth scoring through the reference's
    level sums, and the base range [c0, c1) of each one's windows covering a
    differing base (windows w0..w1: bases w0 .. w1 + l - 1)."""
    W = lengths - l + 1
    w0 = np.maximum(lo - l + 1, 0)
    w1 = np.minimum(hi, W - 1)
    nd = w1 - w0 + 1
    ok = (lo >= 0) & (W > 0) & (nd >= 1) & (nd < VARIANT_ALIAS_FRACTION * W)
    return ok, w0, w1 + l


# ISM sweep (gkm_ism): window pairs of (query, SV) per second, measured on the
# -t 5 configuration (1 kb query against 50,000 SVs of 500 bp: 47.6 ms)
SWEEP_PAIRS_PER_S = 1.0e12
ISM_MAX_BYTES = 16 << 30  # workspace of one query's ISM through the index


def index_ism_ok(kp: KernelParams, idx) -> bool:
    """Whether gkm_index_ism takes this index: bucket records, 8- or 16-bit
    counts, and rows of l windows within what the counts hold."""
    if idx is None or idx.records is None:
        return False
    if kp.weighted:  # 16-bit weighted counts whose 32-bit fallback must not overflow
        return idx.pack == 2 and kp.l * idx.max_weight**2 * max(1, idx.max_sv_windows) < 2**32
    return idx.pack16 and kp.l * max(1, idx.max_sv_windows) < 65536


def ism_use_index(sv: "SVModel", x_lengths, mode="auto") -> bool:
    """Whether ISM of rows of these lengths goes through the l-mer index (3 l + 1
    probed windows per query window: the window itself, and the windows
    covering each position with each of 3 substitutions) rather than the sweep: forced
    by mode True/False where the index applies, else the cheaper by the cost
    model."""
    if mode is False or sv.kp.l > INDEX_MAX_L or sv.kp.dmax < 0:
        return False
    if not index_ism_ok(sv.kp, sv.index()):
        return False
    lmax = int(np.max(x_lengths)) if len(x_lengths) else 0
    if _lib.load().gkm_index_ism_bytes(lmax, sv.kp.l, sv.S, sv.index().n_tiles) > ISM_MAX_BYTES:
        return False  # the per-window and per-position SV parts of one query
    if mode is True:
        return True
    wx = float(np.clip(np.asarray(x_lengths) - sv.kp.l + 1, 0, None).sum())
    index_s, _ = sv._path_seconds((3.0 * sv.kp.l + 1.0) * wx)
    norient = 2 if sv.kp.rc else 1
    ws = float(np.clip(sv.seqs.lengths - sv.kp.l + 1, 0, None).sum())
    return index_s < wx * norient * ws / SWEEP_PAIRS_PER_S


def index_ism(sv: "SVModel", xs: DeviceSeqs, b: int, dx, dmut, out: torch.Tensor,
              stream=None) -> torch.Tensor:
    """out [4, lmax] (contiguous) = the ISM deltas of row b through the l-mer
    index (gkm_index_ism); dx [B] the rows' and dmut [4, lmax, B] the
    mutants' self-kernels (None without normalisation).  Entries of the
    reference base and past the row's length are left as they are."""
    kp, idx = sv.kp, sv.index()
    lx = int(xs.lengths[b])
    fwd, _ = xs.windows(kp.l, stream)
    wt, _ = kp.wtab(xs.lmax)
    ws = torch.empty(max(1, _lib.load().gkm_index_ism_bytes(lx, kp.l, sv.S, idx.n_tiles)),
                     dtype=torch.uint8, device=kp.device)
    wx = max(0, lx - kp.l + 1)
    norm = kp.normalize
    _timed("k_index_ism", lambda: index_bytes(kp, idx, [(3 * kp.l + 1) * wx + kp.l - 1]),
           stream, lambda: _lib.call(
               "gkm_index_ism", _p(fwd[b]), _p(xs.codes[b]), lx, xs.lmax, _p(idx.offsets),
               _p(idx.ids), _p(idx.vals), _p(idx.records), idx.record_tables, _p(idx.masks),
               _p(idx.levels), idx.n_masks, kp.l, kp.nm, _p(wt), sv.S, idx.tile, idx.n_tiles,
               {1: 0, 2: 1, 4: 2}[idx.pack], _p(kp.tab), _p(sv.coef), _p(sv.ds),
               dx.data_ptr() + 8 * b if norm else None,
               dmut.data_ptr() + 8 * b if norm else None, xs.n, int(norm), int(kp.rbf),
               float(kp.gamma), n_cus(kp.device), _p(ws), _p(out), stream_handle(stream)))
    return out


def rows_per_chunk(S: int, nm: int, cap: int = MAX_HIST_BYTES) -> int:
    return max(1, cap // max(1, S * nm * 8))


def sv_hist(sv: SVModel, xs: DeviceSeqs, r0: int, r1: int, *, out=None, stream=None,
            shape=None, use_index="auto") -> torch.Tensor:
    """[r1 - r0, S, nm] histograms of rows r0..r1 against the SVs, through the
    l-mer index or the brute-force kernel (identical integers either way)."""
    if sv.use_index_for(xs.lengths[r0:r1], use_index):
        return index_hist(sv.kp, sv.index(), xs, r0, r1, sv.S, out=out, stream=stream)
    return cross_hist(sv.kp, xs, r0, r1, sv.seqs, sv.groups, out=out, stream=stream, shape=shape)


def table_rows(kp: KernelParams, xs: DeviceSeqs, table: torch.Tensor, bias: float,
               out: torch.Tensor, *, normalize: bool | None = None, stream=None):
    """out[b] = sum_i wt_i table[x_b,i] / sqrt(K_raw(x_b, x_b)) + bias through
    the exact l-mer table (svm.py:559-582); normalize=False gives the raw
    window sums (bias then 0 in the callers that need them)."""
    norm = kp.normalize if normalize is None else normalize
    dx = self_kraw(kp, xs, stream) if norm else None
    fwd, _ = xs.windows(kp.l, stream)
    wt, _ = kp.wtab(xs.lmax)
    _timed("k_table_score", lambda: float(np.clip(xs.lengths - kp.l + 1, 0, None).sum()) * 12.0,
           stream, lambda: _lib.call(
               "gkm_table_score", _p(fwd), _p(xs.len), xs.n, fwd.shape[1], kp.l, _p(wt),
               _p(table), _p(dx), int(norm), float(bias), _p(out), stream_handle(stream)))
    return out


def pairs_apply(sv: SVModel) -> bool:
    """Whether score_rows can take allele pairs through their references'
    level sums (the fused index scoring: no RBF, no position weights)."""
    return sv.wfix is not None and not sv.kp.weighted


def score_rows(sv: SVModel, xs: DeviceSeqs, bias: float, out: torch.Tensor, stream=None,
               shape=None, use_index="auto", table: torch.Tensor | None = None,
               pairs: int = 0):
    """out[b] = sum_s coef_s F(K(x_b, s)) + bias for every row of xs (through
    the exact l-mer table when one is given).  pairs: the last 4 * pairs rows
    are allele pairs in four blocks (references, each alternate's windows
    over its variant, the references' windows there, the alternates); where
    pairs_apply and the index is taken, the alternates are scored from the
    other blocks' level sums instead of probed, with the same bits."""
    kp = sv.kp
    if table is not None:
        return table_rows(kp, xs, table, bias, out, stream=stream)
    dx = self_kraw(kp, xs, stream) if kp.normalize else None
    if sv.wfix is not None and sv.use_index_for(xs.lengths, use_index):
        # one persistent launch: balanced over rows of any lengths (a 103-row
        # serving batch 9.8 -> 7.9 ms against the grid of index_score)
        return index_score_queue(sv, xs, dx, bias, out, stream=stream,
                                 pairs=pairs if pairs_apply(sv) else 0)
    step = rows_per_chunk(sv.S, kp.nm)
    hist = None
    for r0 in range(0, xs.n, step):
        r1 = min(xs.n, r0 + step)
        if hist is None or hist.shape[0] != r1 - r0:
            hist = torch.empty((r1 - r0, sv.S, kp.nm), dtype=torch.int64, device=kp.device)
        sv_hist(sv, xs, r0, r1, out=hist, stream=stream, shape=shape, use_index=use_index)
        if sv.wfix is not None:  # exact level sums, the same bits as index_score
            _lib.call("gkm_combine_fixed", _p(hist), r1 - r0, sv.S, kp.nm, _p(kp.tab),
                      _p(dx[r0:r1]) if dx is not None else None, _p(sv.wfix), sv.fix_shift,
                      int(kp.normalize), float(bias), _p(out[r0:r1]), stream_handle(stream))
        else:
            _lib.call("gkm_combine_score", _p(hist), r1 - r0, sv.S, kp.nm, _p(kp.tab),
                      _p(dx[r0:r1]) if dx is not None else None, _p(sv.ds), _p(sv.coef),
                      int(kp.normalize), int(kp.rbf), kp.gamma, float(bias), _p(out[r0:r1]),
                      stream_handle(stream))
    return out
