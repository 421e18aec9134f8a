"""Goldilocks NTT and K12 Merkle commit on MI355X (gfx950).

Python mirror of the reference's entry points over the C ABI in
include/goldilocks_gpu.h (libgoldilocks_amd.so, loaded with ctypes):

  ntt crate (ntt/src/ntt/mod.rs)    ntt(values), intt(values), strategy(n),
                                    clear_cache(), is_divisor(n)
  pcs crate (pcs/src/k12/mod.rs)    merkle_tree(data), process_layer(input),
                                    k12(data, H), next_layer_size(n)
  composed                          commit(matrix, blowup), ntt_batch_mixed(list)
  binary tree of row hashes         row_tree(matrix), commit_rows(matrix, blowup),
  (no reference function)           row_open(...), row_verify(...)

`values` is a numpy uint64 array (host: copied to HBM and back) or a torch
tensor on the GPU (int64/uint64 storage, transformed in place on the current
stream).  Field words are the raw u64 words of the reference's `[Field]`.
Where the reference panics these functions raise ValueError; device errors
raise RuntimeError.  There is no CPU fallback: if the HIP library is missing
or no gfx950 device is present, calls raise.
"""
import ctypes
import os

import numpy as np

__all__ = [
    "P", "ntt", "intt", "strategy", "Strategy", "clear_cache", "is_divisor", "next_size",
    "merkle_tree", "process_layer", "k12", "next_layer_size", "merkle_tree_size",
    "commit", "ntt_batch_mixed", "fill_random", "lib", "GoldilocksError",
    "commit_rows", "row_tree", "row_tree_size", "row_tree_depth", "row_layers", "row_open", "row_verify",
]

P = 0xFFFFFFFF00000001
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgoldilocks_amd.so")

GL_OK, GL_E_SIZE, GL_E_NOMEM, GL_E_DEVICE, GL_E_ARG, GL_E_UNSUPPORTED, GL_E_BUSY = range(7)


class GoldilocksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing; run `python -m goldilocks_amd.build`")
        # torch ships its own libamdhip64 (same SONAME).  Loading torch first
        # lets this library bind to that copy, so the process has one HIP
        # runtime; loading ours first would make torch load a second one.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        u64, sz, i32, vp = ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p
        psz = ctypes.POINTER(sz)
        sigs = {
            "gl_last_error": (ctypes.c_char_p, []),
            "gl_init": (i32, [i32]),
            "gl_is_divisor": (i32, [u64]),
            "gl_next_size": (u64, [u64]),
            "gl_ntt": (i32, [vp, sz]),
            "gl_intt": (i32, [vp, sz]),
            "gl_ntt_batch": (i32, [vp, sz, sz, i32]),
            "gl_ntt_batch_mixed": (i32, [vp, vp, sz, i32]),
            "gl_ntt_dev": (i32, [vp, sz, sz, i32, vp]),
            "gl_ntt_dev_noalloc": (i32, [vp, sz, sz, i32, vp]),
            "gl_ntt_cols_dev": (i32, [vp, sz, sz, sz, i32, vp]),
            "gl_ntt_mixed_dev": (i32, [vp, vp, sz, i32, vp]),
            "gl_ntt_mixed_dev_unbatched": (i32, [vp, vp, sz, i32, vp]),
            "gl_fill_random_dev": (i32, [vp, sz, u64, vp]),
            "gl_clear_cache": (None, []),
            "gl_plan_passes": (i32, [u64, vp, vp, vp, sz, psz]),
            "gl_variants": (sz, [vp, vp, vp, sz]),
            "gl_plan_set": (i32, [u64, vp, sz]),
            "gl_plan_variants": (i32, [u64, vp, sz, psz]),
            "gl_plan_variants_cols": (i32, [u64, vp, sz, psz]),
            "gl_plan_set_cols": (i32, [u64, vp, sz]),
            "gl_variant_name": (ctypes.c_char_p, [ctypes.c_uint32]),
            "gl_set_twiddle_full_max": (u64, [u64]),
            "gl_timing_enable": (None, [i32]),
            "gl_set_batch_chunk_bytes": (u64, [u64]),
            "gl_timing_pending": (sz, []),
            "gl_timing_collect": (i32, [vp, vp, sz, psz]),
            "gl_batch_plan_create": (i32, [vp, vp, sz, i32, vp]),
            "gl_batch_plan_launch": (i32, [vp, vp]),
            "gl_batch_plan_destroy": (None, [vp]),
            "gl_server_create": (vp, [sz, i32]),
            "gl_server_submit": (i32, [vp, vp, ctypes.POINTER(u64)]),
            "gl_server_poll": (i32, [vp, u64, ctypes.POINTER(i32)]),
            "gl_server_wait": (i32, [vp, u64, ctypes.POINTER(ctypes.c_double)]),
            "gl_server_drain": (i32, [vp]),
            "gl_server_stats_get": (i32, [vp, vp]),
            "gl_server_destroy": (None, [vp]),
            "gl_pow_solve": (i32, [vp, u64, u64, ctypes.POINTER(u64)]),
            "gl_convolve": (i32, [vp, vp, sz]),
            "gl_convolve_dev": (i32, [vp, vp, sz, vp]),
            "gl_k12": (i32, [vp, sz, vp, sz]),
            "gl_next_layer_size": (sz, [sz]),
            "gl_merkle_tree_size": (sz, [sz]),
            "gl_k12_layer": (i32, [vp, sz, vp]),
            "gl_k12_layer_dev": (i32, [vp, sz, vp, vp]),
            "gl_merkle_tree": (i32, [vp, sz, vp, sz, psz]),
            "gl_merkle_tree_dev": (i32, [vp, sz, vp, sz, psz, vp]),
            "gl_merkle_depth": (sz, [sz]),
            "gl_merkle_open_dev": (i32, [vp, sz, vp, vp, sz, vp, vp]),
            "gl_merkle_verify_dev": (i32, [vp, sz, vp, sz, vp, vp, vp]),
            "gl_commit": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, psz]),
            "gl_commit_dev": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, psz, vp]),
            "gl_commit_rows": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, psz]),
            "gl_commit_rows_dev": (i32, [vp, sz, sz, ctypes.c_uint, vp, vp, sz, psz, vp]),
            "gl_row_tree_size": (sz, [sz]),
            "gl_row_tree_depth": (sz, [sz]),
            "gl_row_tree": (i32, [vp, sz, sz, vp, sz, psz]),
            "gl_row_tree_dev": (i32, [vp, sz, sz, vp, sz, psz, vp]),
            "gl_row_open_dev": (i32, [vp, sz, vp, sz, vp, vp]),
            "gl_row_verify_dev": (i32, [vp, sz, sz, vp, sz, vp, vp, vp, vp]),
            "gl_debug_fail_allocs": (None, [i32]),
            "gl_debug_fail_allocs_after": (None, [i32, i32]),
            "gl_transpose": (i32, [vp, sz, sz]),
            "gl_transpose_dev": (i32, [vp, sz, sz, vp]),
            "gl_transpose_oop_dev": (i32, [vp, vp, sz, sz, vp]),
            "gl_bit_reverse": (i32, [vp, sz]),
            "gl_bit_reverse_dev": (i32, [vp, sz, vp]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc == GL_OK:
        return
    msg = lib().gl_last_error().decode(errors="replace")
    if rc in (GL_E_SIZE, GL_E_ARG):
        raise ValueError(msg)
    raise GoldilocksError(rc, msg)


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _stream_ptr(t):
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _host_u64(values):
    if not isinstance(values, np.ndarray) or values.dtype != np.uint64 or not values.flags.c_contiguous:
        raise TypeError("expected a C-contiguous numpy uint64 array")
    return values


def _dev_ptr(t):
    if not t.is_cuda or t.element_size() != 8 or not t.is_contiguous():
        raise TypeError("expected a contiguous 64-bit torch tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def _dev_ptr_arrays(tensors):
    """Device addresses and lengths of a list of GPU tensors as uint64
    arrays, validating each (contiguous, 64-bit, on the GPU) in one pass:
    the batch entry points take thousands of small tensors, where per-tensor
    ctypes objects cost more than the transforms."""
    for t in tensors:
        if not (t.is_cuda and t.element_size() == 8 and t.is_contiguous()):
            raise TypeError("expected contiguous 64-bit torch tensors on the GPU")
    k = len(tensors)
    ptrs = np.fromiter((t.data_ptr() for t in tensors), dtype=np.uint64, count=k)
    ns = np.fromiter((t.numel() for t in tensors), dtype=np.uint64, count=k)
    return ptrs, ns


def _bytes_dev_ptr(t):
    """Device address of a contiguous GPU tensor of any dtype."""
    if not t.is_cuda or not t.is_contiguous():
        raise TypeError("expected a contiguous torch tensor on the GPU")
    return ctypes.c_void_p(t.data_ptr())


def is_divisor(n):
    """ntt/src/divisors.rs:21-23."""
    return bool(lib().gl_is_divisor(n))


def next_size(n):
    """divisors::next (ntt/src/divisors.rs:52-59): smallest supported length >= n."""
    return lib().gl_next_size(n)


def _transform(values, n, count, inverse):
    if _is_torch(values):
        _check(lib().gl_ntt_dev(_dev_ptr(values), n, count, int(inverse), _stream_ptr(values)))
    else:
        v = _host_u64(values)
        _check(lib().gl_ntt_batch(v.ctypes.data_as(ctypes.c_void_p), n, count, int(inverse)))
    return values


def ntt(values):
    """Forward NTT in place (ntt/src/ntt/mod.rs:98-102)."""
    return _transform(values, _len(values), 1, False)


def intt(values):
    """Inverse NTT in place (ntt/src/ntt/mod.rs:104-120)."""
    return _transform(values, _len(values), 1, True)


def _len(values):
    return int(values.numel()) if _is_torch(values) else int(values.size)


class Strategy:
    """strategy(n) (ntt/src/ntt/mod.rs:48-78): an NTT of length n applied to
    every consecutive chunk of n words (trait Ntt, mod.rs:25-31)."""

    def __init__(self, n):
        if not is_divisor(n):
            raise ValueError(f"{n} is not a supported NTT size (does not divide multiplicative order)")
        self.n = n

    def len(self):
        return self.n

    def __len__(self):
        return self.n

    def ntt(self, values, inverse=False):
        total = _len(values)
        if total % self.n:
            raise ValueError("values.len() must be a multiple of the NTT length")
        return _transform(values, self.n, total // self.n, inverse)


def strategy(n):
    return Strategy(n)


def clear_cache():
    """ntt/src/ntt/mod.rs:43-46 (drops device plans and twiddle tables too)."""
    lib().gl_clear_cache()


def plan_info(n):
    """Device schedule for length n: [{"id", "R1", "R2", "R", "T", "kernel"}, ...]."""
    cap = 8
    r1, r2, t = (ctypes.c_uint32 * cap)(), (ctypes.c_uint32 * cap)(), (ctypes.c_uint32 * cap)()
    ids = (ctypes.c_uint32 * cap)()
    npass = ctypes.c_size_t(0)
    _check(lib().gl_plan_passes(n, r1, r2, t, cap, ctypes.byref(npass)))
    _check(lib().gl_plan_variants(n, ids, cap, ctypes.byref(npass)))
    out = []
    for i in range(npass.value):
        name = lib().gl_variant_name(ids[i])
        out.append({"id": ids[i], "R1": r1[i], "R2": r2[i], "R": r1[i] * r2[i], "T": t[i],
                    "kernel": name.decode() if name else "?"})
    return out


def variants():
    """Compiled pass kernels (and the radix-257 Rader kernel, last):
    list of {"id", "R1", "R2", "R", "T", "name"}."""
    n = lib().gl_variants(None, None, None, 0)
    r1, r2, t = (ctypes.c_uint32 * n)(), (ctypes.c_uint32 * n)(), (ctypes.c_uint32 * n)()
    lib().gl_variants(r1, r2, t, n)
    return [{"id": i, "R1": r1[i], "R2": r2[i], "R": r1[i] * r2[i], "T": t[i],
             "name": lib().gl_variant_name(i).decode()} for i in range(n)]


def set_batch_chunk_bytes(nbytes):
    """Chunk size of the grouped batch path (results do not depend on it);
    returns the previous value."""
    return lib().gl_set_batch_chunk_bytes(int(nbytes))


def timing_enable(on=True):
    """Bracket every NTT pass launch with HIP events on its stream."""
    lib().gl_timing_enable(int(bool(on)))


def timing_collect():
    """Wait for the passes recorded since the last call; returns a list of
    (kernel name, milliseconds) in launch order."""
    cap = lib().gl_timing_pending() + 64
    ids, ms = (ctypes.c_uint32 * cap)(), (ctypes.c_float * cap)()
    cnt = ctypes.c_size_t(0)
    _check(lib().gl_timing_collect(ids, ms, cap, ctypes.byref(cnt)))
    out = []
    for i in range(min(cnt.value, cap)):
        name = lib().gl_variant_name(ids[i])
        out.append((name.decode() if name else "?", ms[i]))
    return out


def set_twiddle_full_max(words):
    """Threshold (sub-problem length) up to which a pass's output twiddles
    come from one full table (up to 4x that for passes whose instance stride
    is >= 32); larger passes use the two-level power table.  Results are
    identical; returns the previous value."""
    return lib().gl_set_twiddle_full_max(int(words))


def plan_set(n, variant_ids, cols=False):
    """Install an explicit schedule for length n (empty list: default);
    cols=True sets the schedule of column transforms of length n."""
    k = len(variant_ids)
    ids = (ctypes.c_uint32 * max(k, 1))(*variant_ids)
    _check((lib().gl_plan_set_cols if cols else lib().gl_plan_set)(n, ids, k))


def plan_ids(n, cols=False):
    """Variant ids of the schedule for length n (column transforms: cols=True)."""
    cap = 8
    ids = (ctypes.c_uint32 * cap)()
    npass = ctypes.c_size_t(0)
    _check((lib().gl_plan_variants_cols if cols else lib().gl_plan_variants)(n, ids, cap, ctypes.byref(npass)))
    return [ids[i] for i in range(npass.value)]


def ntt_cols(matrix, inverse=False):
    """Transform each column of a row-major (rows x cols) GPU tensor in place."""
    rows, cols = matrix.shape
    _check(lib().gl_ntt_cols_dev(_dev_ptr(matrix), rows, cols, 1, int(inverse), _stream_ptr(matrix)))
    return matrix


def ntt_batch_mixed(arrays, inverse=False, batched=True):
    """Transform every array of a list in place (mixed lengths/radices).
    Host numpy arrays or GPU tensors (all of one kind)."""
    k = len(arrays)
    if k == 0:
        return arrays
    if _is_torch(arrays[0]):
        ptrs, ns = _dev_ptr_arrays(arrays)
        fn = lib().gl_ntt_mixed_dev if batched else lib().gl_ntt_mixed_dev_unbatched
        _check(fn(ptrs.ctypes.data, ns.ctypes.data, k, int(inverse), _stream_ptr(arrays[0])))
    else:
        ns = (ctypes.c_size_t * k)(*[_len(a) for a in arrays])
        ptrs = (ctypes.c_void_p * k)(*[_host_u64(a).ctypes.data for a in arrays])
        _check(lib().gl_ntt_batch_mixed(ptrs, ns, k, int(inverse)))
    return arrays


# ---------------------------------------------------------------- permutations
def transpose(values, rows, cols, out=None):
    """permute::transpose (ntt/src/permute/mod.rs:53-58): the flat rows x
    cols row-major matrix becomes its cols x rows transpose, in place (numpy
    array or contiguous GPU tensor).  out (GPU tensor, distinct from values):
    write the transpose there instead and leave values untouched."""
    if _is_torch(values):
        if out is not None:
            _check(lib().gl_transpose_oop_dev(_dev_ptr(values), _dev_ptr(out), rows, cols, _stream_ptr(values)))
            return out
        _check(lib().gl_transpose_dev(_dev_ptr(values), rows, cols, _stream_ptr(values)))
        return values
    if out is not None:
        raise ValueError("out= is for GPU tensors")
    v = _host_u64(values)
    _check(lib().gl_transpose(v.ctypes.data_as(ctypes.c_void_p), rows, cols))
    return values


def bit_reverse(values):
    """permute::bit_reverse (ntt/src/permute/mod.rs:71-74): values[i] <-
    values[rev(i)] in place; the length must be a power of two."""
    n = _len(values)
    if _is_torch(values):
        _check(lib().gl_bit_reverse_dev(_dev_ptr(values), n, _stream_ptr(values)))
    else:
        v = _host_u64(values)
        _check(lib().gl_bit_reverse(v.ctypes.data_as(ctypes.c_void_p), n))
    return values


def fill_random(tensor, seed):
    """Seeded uniform [0, p) words generated on the GPU (same map as the tests)."""
    _check(lib().gl_fill_random_dev(_dev_ptr(tensor), _len(tensor), seed, _stream_ptr(tensor)))
    return tensor


# ---------------------------------------------------------------- K12 / Merkle

def next_layer_size(n):
    """pcs/src/k12/mod.rs:33-35."""
    return lib().gl_next_layer_size(n)


def merkle_tree_size(n):
    return lib().gl_merkle_tree_size(n)


def k12(data: bytes, H: int = 32) -> bytes:
    """k12::<H>(input) (pcs/src/k12/generic.rs:62-84); len(data) <= 166."""
    data = bytes(data)
    out = ctypes.create_string_buffer(max(H, 1))
    _check(lib().gl_k12(data, len(data), out, H))
    return out.raw[:H]


def process_layer(data: bytes) -> bytes:
    """process_layer (pcs/src/k12/generic.rs:86-94): one 32-byte K12 hash per
    160-byte block (last block may be short)."""
    data = bytes(data)
    out = ctypes.create_string_buffer(max(1, next_layer_size(len(data))))
    _check(lib().gl_k12_layer(data, len(data), out))
    return out.raw[: next_layer_size(len(data))]


def _split_layers(raw, n):
    layers, off = [], 0
    while n > 32:
        n = next_layer_size(n)
        layers.append(raw[off:off + n])
        off += n
    return layers


def merkle_tree(data):
    """merkle_tree (pcs/src/k12/mod.rs:41-60): list of layers, root last.

    bytes / numpy input: hashed on the GPU, layers returned as bytes.
    GPU tensor input (any dtype, contiguous): layers returned as one uint8
    GPU tensor of all layers concatenated (split with next_layer_size)."""
    if _is_torch(data):
        import torch
        nbytes = data.numel() * data.element_size()
        need = merkle_tree_size(nbytes)
        out = torch.empty(max(need, 1), dtype=torch.uint8, device=data.device)
        got = ctypes.c_size_t(0)
        _check(lib().gl_merkle_tree_dev(ctypes.c_void_p(data.data_ptr()), nbytes,
                                        ctypes.c_void_p(out.data_ptr()), need, ctypes.byref(got),
                                        _stream_ptr(data)))
        return out[:need]
    if isinstance(data, np.ndarray):
        data = np.ascontiguousarray(data).tobytes()
    data = bytes(data)
    need = merkle_tree_size(len(data))
    out = ctypes.create_string_buffer(max(need, 1))
    got = ctypes.c_size_t(0)
    _check(lib().gl_merkle_tree(data, len(data), out, need, ctypes.byref(got)))
    return _split_layers(out.raw[:need], len(data))


def commit(matrix, blowup=2):
    """Column coset LDE (intt, zero-extend, x GENERATOR^i, ntt) of a row-major
    rows x cols matrix, then merkle_tree over the LDE bytes.

    numpy input -> (lde ndarray, [layer bytes]); GPU tensor input ->
    (lde tensor, layers uint8 tensor)."""
    rows, cols = matrix.shape
    ext = rows * blowup
    need = merkle_tree_size(ext * cols * 8)
    got = ctypes.c_size_t(0)
    if _is_torch(matrix):
        import torch
        lde = torch.empty((ext, cols), dtype=matrix.dtype, device=matrix.device)
        layers = torch.empty(max(need, 1), dtype=torch.uint8, device=matrix.device)
        _check(lib().gl_commit_dev(_dev_ptr(matrix), rows, cols, blowup, _dev_ptr(lde),
                                   ctypes.c_void_p(layers.data_ptr()), need, ctypes.byref(got),
                                   _stream_ptr(matrix)))
        return lde, layers[:need]
    m = _host_u64(np.ascontiguousarray(matrix, dtype=np.uint64))
    lde = np.zeros((ext, cols), dtype=np.uint64)
    out = ctypes.create_string_buffer(max(need, 1))
    _check(lib().gl_commit(m.ctypes.data_as(ctypes.c_void_p), rows, cols, blowup,
                           lde.ctypes.data_as(ctypes.c_void_p), out, need, ctypes.byref(got)))
    return lde, _split_layers(out.raw[:need], ext * cols * 8)


# ---------------------------------------------------------------- row tree
def row_tree_size(leaves):
    """Bytes of every layer of the row tree over `leaves` row hashes."""
    return lib().gl_row_tree_size(leaves)


def row_tree_depth(leaves):
    """Layers above the leaves."""
    return lib().gl_row_tree_depth(leaves)


def row_layers(raw, leaves):
    """Split the concatenated layers (leaves first) into a list, root last."""
    layers, n, off = [], leaves, 0
    while True:
        layers.append(raw[off:off + 32 * n])
        off += 32 * n
        if n <= 1:
            return layers
        n = (n + 1) // 2


def row_tree(matrix):
    """Binary tree of row hashes of a row-major matrix of words (leaf r =
    K12 of row r's bytes, parent = K12(left || right), the last node of an
    odd layer moved up; gl_row_tree).  numpy input -> list of layer bytes,
    leaves first; GPU tensor input -> one uint8 GPU tensor of all layers."""
    nrows, cols = matrix.shape
    need = row_tree_size(nrows)
    got = ctypes.c_size_t(0)
    if _is_torch(matrix):
        import torch
        out = torch.empty(max(need, 8) // 8, dtype=torch.int64, device=matrix.device).view(torch.uint8)
        _check(lib().gl_row_tree_dev(_dev_ptr(matrix), nrows, cols, ctypes.c_void_p(out.data_ptr()), need,
                                     ctypes.byref(got), _stream_ptr(matrix)))
        return out[:need]
    m = _host_u64(np.ascontiguousarray(matrix, dtype=np.uint64))
    out = ctypes.create_string_buffer(max(need, 1))
    _check(lib().gl_row_tree(m.ctypes.data_as(ctypes.c_void_p), nrows, cols, out, need, ctypes.byref(got)))
    return row_layers(out.raw[:need], nrows)


def commit_rows(matrix, blowup=2):
    """commit's column coset extension with the row tree over its rows.

    numpy input -> (lde ndarray, [layer bytes], leaves first); GPU tensor
    input -> (lde tensor, layers uint8 tensor)."""
    rows, cols = matrix.shape
    ext = rows * blowup
    need = row_tree_size(ext)
    got = ctypes.c_size_t(0)
    if _is_torch(matrix):
        import torch
        lde = torch.empty((ext, cols), dtype=matrix.dtype, device=matrix.device)
        layers = torch.empty(max(need, 8) // 8, dtype=torch.int64, device=matrix.device).view(torch.uint8)
        _check(lib().gl_commit_rows_dev(_dev_ptr(matrix), rows, cols, blowup, _dev_ptr(lde),
                                        ctypes.c_void_p(layers.data_ptr()), need, ctypes.byref(got),
                                        _stream_ptr(matrix)))
        return lde, layers[:need]
    m = _host_u64(np.ascontiguousarray(matrix, dtype=np.uint64))
    lde = np.zeros((ext, cols), dtype=np.uint64)
    out = ctypes.create_string_buffer(max(need, 1))
    _check(lib().gl_commit_rows(m.ctypes.data_as(ctypes.c_void_p), rows, cols, blowup,
                                lde.ctypes.data_as(ctypes.c_void_p), out, need, ctypes.byref(got)))
    return lde, row_layers(out.raw[:need], ext)


def row_open(layers, leaves, idx):
    """Openings of leaves `idx` (int64 GPU tensor) from the row tree's layers
    (the uint8 GPU tensor row_tree / commit_rows returned): a uint8 GPU
    tensor (len(idx), depth, 32) of sibling hashes, zeros where a node moves
    up unchanged."""
    import torch
    depth = row_tree_depth(leaves)
    k = idx.numel()
    out = torch.zeros((k, depth * 4), dtype=torch.int64, device=layers.device)
    _check(lib().gl_row_open_dev(_bytes_dev_ptr(layers), leaves, _dev_ptr(idx), k, _dev_ptr(out),
                                 _stream_ptr(layers)))
    return out.view(torch.uint8).view(k, depth, 32)


def row_verify(rows, leaves, idx, paths, root):
    """Check opened rows (k x cols GPU tensor of the rows' words) at leaves
    `idx` against openings (layout of row_open) and a 32-byte root; returns a
    bool GPU tensor, one flag per opening."""
    import torch
    k, cols = rows.shape
    ok = torch.zeros(k, dtype=torch.int32, device=rows.device)
    root8 = torch.empty(4, dtype=torch.int64, device=rows.device)
    root8.view(torch.uint8).copy_(root.reshape(-1)[:32])
    paths8 = paths.contiguous()
    _check(lib().gl_row_verify_dev(_dev_ptr(rows), cols, leaves, _dev_ptr(idx), k, _bytes_dev_ptr(paths8),
                                   _dev_ptr(root8), _bytes_dev_ptr(ok), _stream_ptr(rows)))
    return ok.bool()


# ---------------------------------------------------------------- PoW / convolution

def threshold_from_bits(bits):
    """pcs/src/pow.rs:15-17."""
    return (2**64 - 1) - int(2.0 ** (64.0 - bits))


def pow_verify(challenge: bytes, threshold: int, nonce: int) -> bool:
    """pcs/src/pow.rs:32-37 (one K12 hash on the GPU)."""
    h = k12(bytes(challenge) + int(nonce).to_bytes(8, "little"), 8)
    return int.from_bytes(h, "little") >= threshold


def pow_solve(challenge: bytes, threshold: int, max_nonce=1 << 40) -> int:
    """pcs/src/pow.rs:39-60: the smallest satisfying nonce (GPU scan)."""
    out = ctypes.c_uint64(0)
    _check(lib().gl_pow_solve(bytes(challenge), threshold, max_nonce, ctypes.byref(out)))
    return out.value


def convolve(a, b_ntt):
    """ntt/src/convolve.rs:13-20: a <- intt(ntt(a) * b_ntt), Field (Montgomery)
    pointwise product on raw words.  numpy arrays or GPU tensors, in place."""
    n = _len(a)
    if _len(b_ntt) != n:
        raise ValueError("lengths differ")
    if _is_torch(a):
        _check(lib().gl_convolve_dev(_dev_ptr(a), _dev_ptr(b_ntt), n, _stream_ptr(a)))
    else:
        _check(lib().gl_convolve(_host_u64(a).ctypes.data_as(ctypes.c_void_p),
                                 _host_u64(np.ascontiguousarray(b_ntt, dtype=np.uint64)).ctypes.data_as(ctypes.c_void_p), n))
    return a


# ---------------------------------------------------------------- batch plans

class BatchPlan:
    """A mixed-size batch over fixed GPU tensors, planned once and replayed
    as one captured hipGraph launch (gl_batch_plan_*)."""

    def __init__(self, tensors, inverse=False):
        k = len(tensors)
        self._tensors = list(tensors)  # keep the buffers alive
        ptrs, ns = _dev_ptr_arrays(self._tensors)
        h = ctypes.c_void_p()
        _check(lib().gl_batch_plan_create(ptrs.ctypes.data, ns.ctypes.data, k, int(inverse), ctypes.byref(h)))
        self._h = h

    def launch(self, stream=None):
        import torch
        s = stream if stream is not None else torch.cuda.current_stream()
        _check(lib().gl_batch_plan_launch(self._h, ctypes.c_void_p(s.cuda_stream)))

    def close(self):
        if self._h:
            lib().gl_batch_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def merkle_depth(leaf_len):
    """Number of hashed layers of a tree over leaf_len bytes."""
    return lib().gl_merkle_depth(leaf_len)


def merkle_open(leaves, layers, blocks):
    """Openings of leaf blocks `blocks` (int64 GPU tensor of 160-byte block
    indices) from a device-resident tree: `leaves` (any contiguous GPU
    tensor, its bytes are the leaves) and `layers` (the uint8 tensor
    merkle_tree returned for it).  Returns a uint8 GPU tensor of shape
    (len(blocks), depth, 160): for each layer l, block blocks[i] // 5**l of
    layer l, zero padded."""
    import torch
    leaf_len = leaves.numel() * leaves.element_size()
    depth = merkle_depth(leaf_len)
    k = blocks.numel()
    out = torch.zeros((k, depth, 160), dtype=torch.uint8, device=leaves.device)
    _check(lib().gl_merkle_open_dev(_bytes_dev_ptr(leaves), leaf_len, _bytes_dev_ptr(layers), _dev_ptr(blocks), k,
                                    _bytes_dev_ptr(out), _stream_ptr(leaves)))
    return out


def merkle_verify(paths, leaf_len, blocks, root):
    """Check openings (layout of merkle_open) against a 32-byte root (uint8
    GPU tensor); returns a bool GPU tensor, one flag per opening."""
    import torch
    k = blocks.numel()
    ok = torch.zeros(k, dtype=torch.int32, device=paths.device)
    _check(lib().gl_merkle_verify_dev(_bytes_dev_ptr(paths), leaf_len, _dev_ptr(blocks), k, _bytes_dev_ptr(root),
                                      _bytes_dev_ptr(ok), _stream_ptr(paths)))
    return ok.bool()


# ---------------------------------------------------------------- serving

GL_JOB_NTT, GL_JOB_INTT, GL_JOB_COMMIT, GL_JOB_COMMIT_ROWS = 0, 1, 2, 3


class _Job(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("device", ctypes.c_int), ("data", ctypes.c_void_p),
                ("n", ctypes.c_size_t), ("count", ctypes.c_size_t), ("blowup", ctypes.c_uint),
                ("lde_out", ctypes.c_void_p), ("layers_out", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("rejected", ctypes.c_uint64), ("p50_ms", ctypes.c_double), ("p99_ms", ctypes.c_double),
                ("mean_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("elements", ctypes.c_double),
                ("busy_seconds", ctypes.c_double), ("peak_bytes", ctypes.c_size_t)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class Server:
    """Continuous-batching NTT / commit server (gl_server_*).  Jobs are
    admitted as streams and HBM budget free up; their work overlaps on the
    device.  Buffers (numpy for host jobs, GPU tensors for device jobs) must
    stay alive until the job is done."""

    def __init__(self, hbm_budget_bytes=0, slots=8):
        # 8 slots: at saturation the mixed serving workload runs 42 / 62 /
        # 73 / 77 Gelem/s with 1 / 2 / 4 / 8 (16 adds nothing)
        h = lib().gl_server_create(hbm_budget_bytes, slots)
        if not h:
            raise GoldilocksError(GL_E_DEVICE, lib().gl_last_error().decode())
        self._h = ctypes.c_void_p(h)
        self._keep = {}

    def _submit(self, job, keep):
        t = ctypes.c_uint64(0)
        _check(lib().gl_server_submit(self._h, ctypes.byref(job), ctypes.byref(t)))
        self._keep[t.value] = keep
        return t.value

    def submit_ntt(self, values, n=None, inverse=False):
        """Transform `values` (numpy uint64 or GPU tensor) in place as
        len(values) // n transforms of length n."""
        n = n or _len(values)
        dev = _is_torch(values)
        ptr = _dev_ptr(values).value if dev else _host_u64(values).ctypes.data
        job = _Job(GL_JOB_INTT if inverse else GL_JOB_NTT, int(dev), ptr, n, _len(values) // n, 0, None, None)
        return self._submit(job, values)

    def submit_commit(self, matrix, blowup, lde_out, layers_out, row_tree=False):
        """Commit a rows x cols matrix; lde_out ((blowup*rows) x cols) and
        layers_out (merkle_tree_size bytes, or row_tree_size(blowup*rows)
        with row_tree=True: the commit_rows form) receive the results."""
        rows, cols = matrix.shape
        dev = _is_torch(matrix)
        if dev:
            ptrs = (_dev_ptr(matrix).value, _dev_ptr(lde_out).value, ctypes.c_void_p(layers_out.data_ptr()).value)
        else:
            ptrs = (_host_u64(matrix).ctypes.data, _host_u64(lde_out).ctypes.data, layers_out.ctypes.data)
        job = _Job(GL_JOB_COMMIT_ROWS if row_tree else GL_JOB_COMMIT, int(dev), ptrs[0], rows, cols, blowup,
                   ptrs[1], ptrs[2])
        return self._submit(job, (matrix, lde_out, layers_out))

    def poll(self, ticket):
        done = ctypes.c_int(0)
        _check(lib().gl_server_poll(self._h, ticket, ctypes.byref(done)))
        return bool(done.value)

    def wait(self, ticket):
        """Block until the job is done; returns its submit-to-done latency in ms."""
        lat = ctypes.c_double(0)
        rc = lib().gl_server_wait(self._h, ticket, ctypes.byref(lat))
        self._keep.pop(ticket, None)
        _check(rc)
        return lat.value

    def drain(self):
        _check(lib().gl_server_drain(self._h))

    def stats(self):
        st = ServerStats()
        _check(lib().gl_server_stats_get(self._h, ctypes.byref(st)))
        return st.as_dict()

    def close(self):
        if self._h:
            lib().gl_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
