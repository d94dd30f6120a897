"""Exact kNN with a resident reference set (BF16 MFMA prefilter + FP32 rerank).

Mirrors ``exact_neighbors(query, reference, k, exclude_self)``
(crates/umap/src/standard_umap/neighbors.rs:17-77): indices [q, k] (int32)
and distances [q, k] (float32, square roots), ties to the lower reference
index, sentinel id n / distance sqrt(f32::MAX) for unfilled slots.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class KnnIndex:
    def __init__(self, reference, max_queries: int = 4096):
        import torch

        if isinstance(reference, torch.Tensor):
            ref = reference.detach().to(dtype=torch.float32).contiguous()
            on_device = ref.is_cuda
            ptr = ref.data_ptr() if on_device else ref.numpy().ctypes.data
            keep = ref
        else:
            keep = np.ascontiguousarray(reference, np.float32)
            on_device = False
            ptr = keep.ctypes.data
        self.n, self.d = keep.shape
        self.max_queries = int(max_queries)
        handle = C.c_void_p()
        _lib.check(_lib.lib().umaphip_knn_index_create(ptr, int(on_device), self.n, self.d,
                                                       self.max_queries, C.byref(handle)))
        self._handle = handle

    def query_device(self, queries, k: int, *, exclude_self=False, query_offset=0):
        """queries: torch f32 [q, d] on the device; returns device tensors."""
        import torch

        q = queries.contiguous()
        if q.dtype != torch.float32 or not q.is_cuda or q.shape[1] != self.d:
            raise ValueError("queries must be a float32 [q, d] tensor on the HIP device")
        idx = torch.empty((q.shape[0], k), dtype=torch.int32, device=q.device)
        dist = torch.empty((q.shape[0], k), dtype=torch.float32, device=q.device)
        _lib.check(_lib.lib().umaphip_knn_query(
            self._handle, q.data_ptr(), q.shape[0], int(k), int(bool(exclude_self)),
            int(query_offset), idx.data_ptr(), dist.data_ptr(),
            torch.cuda.current_stream().cuda_stream))
        return idx, dist

    def candidates_device(self, queries):
        """The BF16 prefilter's merged candidate lists before the exact rerank:
        (ids [q, 64] best first, counts [q]) as device tensors."""
        import torch

        q = queries.contiguous()
        if q.dtype != torch.float32 or not q.is_cuda or q.shape[1] != self.d:
            raise ValueError("queries must be a float32 [q, d] tensor on the HIP device")
        ids = torch.full((q.shape[0], 64), -1, dtype=torch.int32, device=q.device)
        counts = torch.empty(q.shape[0], dtype=torch.int32, device=q.device)
        _lib.check(_lib.lib().umaphip_knn_candidates(
            self._handle, q.data_ptr(), q.shape[0], ids.data_ptr(), counts.data_ptr(),
            torch.cuda.current_stream().cuda_stream))
        return ids, counts

    def query(self, queries, k: int, *, exclude_self=False, query_offset=0, batch=None):
        """Host arrays in and out; batches of at most max_queries rows."""
        import torch

        q = np.ascontiguousarray(queries, np.float32)
        batch = batch or self.max_queries
        out_i = np.empty((q.shape[0], k), np.int32)
        out_d = np.empty((q.shape[0], k), np.float32)
        for s in range(0, q.shape[0], batch):
            qt = torch.from_numpy(q[s:s + batch]).cuda()
            i, d = self.query_device(qt, k, exclude_self=exclude_self,
                                     query_offset=query_offset + s)
            out_i[s:s + batch] = i.cpu().numpy()
            out_d[s:s + batch] = d.cpu().numpy()
        return out_i, out_d

    def set_rescan(self, every_query: bool) -> None:
        """Rescan every query exactly in FP32 (tests); False restores the
        per-query certificate, which rescans only the queries it leaves open."""
        _lib.check(_lib.lib().umaphip_knn_set_rescan(self._handle, int(bool(every_query))))

    def certificate_stats(self):
        """(queries of the last query call rescanned exactly, largest BF16
        rounding of a reference row relative to its norm)."""
        import torch

        n = C.c_int64()
        eps = C.c_float()
        _lib.check(_lib.lib().umaphip_knn_certificate_stats(
            self._handle, torch.cuda.current_stream().cuda_stream, C.byref(n), C.byref(eps)))
        return int(n.value), float(eps.value)

    def close(self):
        if getattr(self, "_handle", None):
            _lib.lib().umaphip_knn_index_destroy(self._handle)
            self._handle = None

    def __del__(self):
        # at interpreter shutdown the binding may already be torn down
        try:
            self.close()
        except (TypeError, AttributeError):
            pass
