"""Batched form: many independent products C_i = A_i * B_i as one
block-diagonal product, planned once and replayed as one captured hipGraph
(ams_batched_plan / ams_batched_capture / ams_graph_launch of the C ABI).
The reference has no batched form.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Tuple

import numpy as np
import torch

from . import _lib as L
from .csr import Csr
from .device import Context, DeviceCsr


def stack(problems: List[Tuple[Csr, Csr]]):
    """Block-diagonal stacking.  Returns (A_stack, B_stack, a_row_off,
    b_row_off, c_col_off) with A_i's columns shifted to B_i's rows and B_i's
    columns shifted to C_i's output columns."""
    a_row = [0]
    b_row = [0]
    c_col = [0]
    for a, b in problems:
        if a.num_col != b.num_row:
            raise ValueError("A_i columns must equal B_i rows")
        a_row.append(a_row[-1] + a.num_row)
        b_row.append(b_row[-1] + b.num_row)
        c_col.append(c_col[-1] + b.num_col)

    def cat(mats, col_shift):
        rps, cols, vals = [np.zeros(1, np.int64)], [], []
        base = 0
        for (m, sh) in zip(mats, col_shift):
            rps.append(np.asarray(m.row_ptrs[1:], np.int64) + base)
            base += m.num_nnz
            cols.append(np.asarray(m.cols, np.int64) + sh)
            vals.append(m.rats)
        return np.concatenate(rps), np.concatenate(cols).astype(np.int32), np.concatenate(vals)

    rp, cols, vals = cat([p[0] for p in problems], b_row[:-1])
    A = Csr(a_row[-1], b_row[-1], rp, cols, vals)
    rp, cols, vals = cat([p[1] for p in problems], c_col[:-1])
    B = Csr(b_row[-1], c_col[-1], rp, cols, vals)
    return A, B, np.array(a_row, np.int64), np.array(b_row, np.int64), np.array(c_col, np.int64)


class Stacked:
    """The block-diagonal stack of a batch, uploaded once (reused by every
    plan of the autotuner)."""

    def __init__(self, problems: List[Tuple[Csr, Csr]]):
        A, B, self.a_off, self.b_off, self.c_off = stack(problems)
        self.A = DeviceCsr.from_host(A)
        self.B = DeviceCsr.from_host(B)
        torch.cuda.synchronize()  # uploads ran on the current stream; plans run on the context's


class BatchedProduct:
    """Plans a batch once; ``run()`` replays one captured graph."""

    def __init__(self, problems: List[Tuple[Csr, Csr]], ams: Context | None = None, capture: bool = True,
                 stacked: "Stacked | None" = None):
        self.problems = problems
        self.stream = torch.cuda.Stream() if ams is None else ams.stream
        self.ams = ams or Context(stream=self.stream)
        stacked = stacked or Stacked(problems)
        self.a_off, self.b_off, self.c_off = stacked.a_off, stacked.b_off, stacked.c_off
        self.A, self.B = stacked.A, stacked.B
        A = stacked.A
        lib = L.lib()
        sa, sb = self.A.c_struct(), self.B.c_struct()
        cap = C.c_int64()
        ptr = lambda x: x.ctypes.data_as(C.c_void_p)  # noqa: E731
        self.ams._chk(lib.ams_batched_plan(self.ams.handle, C.byref(sa), C.byref(sb), self.A.dtype_code,
                                           len(problems), ptr(self.a_off), ptr(self.b_off), ptr(self.c_off),
                                           C.byref(cap)), "ams_batched_plan")
        self.cap = cap.value
        dev = self.A.val.device
        self.c_rp = torch.empty(A.num_row + 1, dtype=torch.int64, device=dev)
        self.c_col = torch.empty(max(self.cap, 1), dtype=torch.int32, device=dev)
        self.c_val = torch.empty(max(self.cap, 1), dtype=self.A.val.dtype, device=dev)
        self.nnz = torch.zeros(1, dtype=torch.int64, device=dev)
        self._sa, self._sb = sa, sb
        self.graph = None
        if capture:
            g = C.c_void_p()
            self.ams._chk(lib.ams_batched_capture(self.ams.handle, C.byref(sa), C.byref(sb), self.A.dtype_code,
                                                  self.c_rp.data_ptr(), self.c_col.data_ptr(), self.c_val.data_ptr(),
                                                  self.cap, self.nnz.data_ptr(), C.byref(g)),
                          "ams_batched_capture")
            self.graph = g

    def run(self):
        lib = L.lib()
        if self.graph is not None:
            self.ams._chk(lib.ams_graph_launch(self.ams.handle, self.graph), "ams_graph_launch")
        else:
            self.ams._chk(lib.ams_batched_run(self.ams.handle, C.byref(self._sa), C.byref(self._sb),
                                              self.A.dtype_code, self.c_rp.data_ptr(), self.c_col.data_ptr(),
                                              self.c_val.data_ptr(), self.cap, self.nnz.data_ptr()),
                          "ams_batched_run")

    def synchronize(self):
        if self.stream is None:
            torch.cuda.synchronize()
        else:
            self.stream.synchronize()

    def check(self):
        """Raises when the kernels of the last run raised error flags."""
        flags = self.ams.error_flags()
        if flags:
            raise L.AmsError(f"batched run raised kernel error flags 0x{flags:x}")

    def results(self) -> List[Csr]:
        """Per-problem products (host copies, columns rebased) of the last
        run, after checking its kernel error flags."""
        self.synchronize()
        self.check()
        rp = self.c_rp.cpu().numpy()
        nnz = int(rp[-1])
        cols = self.c_col[:nnz].cpu().numpy().astype(np.int64)
        vals = self.c_val[:nnz].cpu().numpy()
        out = []
        for i, (a, b) in enumerate(self.problems):
            r0, r1 = self.a_off[i], self.a_off[i + 1]
            s, e = rp[r0], rp[r1]
            out.append(Csr(a.num_row, b.num_col, rp[r0:r1 + 1] - s, (cols[s:e] - self.c_off[i]).astype(np.int32),
                           vals[s:e]))
        return out

    def close(self):
        if self.graph is not None:
            L.lib().ams_graph_destroy(self.graph)
            self.graph = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# Coordinate search of the batched autotuner: each parameter in turn over
# its candidates, keeping the best so far (the same space as ams_autotune).
TUNE_SPACE = [
    ("hash", [(682, 682), (1365, 1365), (2730, 2730), (1365, 2730), (2730, 5461), (5461, 5461)]),
    ("dense_width", [4288, 8448, 12800, 17344]),
    ("acc_thres", [0.0, 1.25, 2.0, 4.0]),
    ("bitmap_ratio", [0.25, 1.0, 4.0]),
    ("dense_density", [0.02, 0.05, 0.2, 1.0]),
    ("esc_sort", [0, 1, 2, 3, 4]),
]


def autotune_batched(problems: List[Tuple[Csr, Csr]], ams: Context, reps: int = 2, stacked: "Stacked | None" = None):
    """Picks routing thresholds and LDS geometry for a batch by timing the
    planned, unsynchronised batched run (plan outside the timing) over the
    coordinate search in TUNE_SPACE.  Leaves the best parameters set on
    `ams` and returns (params dict, ms per batched run)."""
    stacked = stacked or Stacked(problems)

    def timed(params):
        ams.set_params(**params)
        bp = BatchedProduct(problems, ams=ams, capture=False, stacked=stacked)
        bp.run()
        bp.synchronize()
        t0 = torch.cuda.Event(enable_timing=True)
        t1 = torch.cuda.Event(enable_timing=True)
        stream = bp.stream if bp.stream is not None else torch.cuda.current_stream()
        t0.record(stream)
        for _ in range(reps):
            bp.run()
        t1.record(stream)
        bp.synchronize()
        bp.close()
        return t0.elapsed_time(t1) / reps

    best = ams.params()
    best_ms = timed(best)
    for name, cands in TUNE_SPACE:
        for cand in cands:
            p = dict(best)
            if name == "hash":
                p["sym_hash_max"], p["num_hash_max"] = cand
            else:
                p[name] = cand
                if name == "dense_width" and p["bitmap_bits"] < cand:
                    p["bitmap_bits"] = cand
            try:
                ms = timed(p)
            except L.AmsError:
                continue  # outside this batch's limits (LDS, table sizes)
            if ms < best_ms:
                best, best_ms = p, ms
    ams.set_params(**best)
    return best, best_ms
