"""ctypes binding of the TAPP C API.

The same binding drives any library that exports the TAPP API: this
package's libtapp_amd.so (the GPU implementation) and, in the test suite,
the compiled reference library.  It mirrors the ctypes layer the reference
uses in its own Python tests (test/test.py:14-120 of
TAPPorg/reference-implementation): handles are intptr_t, data is passed as
raw pointers, scalars by pointer.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtapp_amd.so")
TRACE_PATH = os.path.join(HERE, "lib", "libtapp_trace.so")

# api/include/tapp/datatype.h:17-95, product.h:17-21
F32, F64, C32, C64, F16, BF16 = 0, 1, 2, 3, 4, 5
DEFAULT_PREC, F32F32_ACCUM_F32, F64F64_ACCUM_F64 = -1, 0, 1
F16F16_ACCUM_F16, F16F16_ACCUM_F32, BF16BF16_ACCUM_F32 = 4, 5, 6
IDENTITY, CONJUGATE = 0, 1

# tapp_amd.h
ATTR_SPLIT_BF16 = 0x1001
ATTR_KERNEL = 0x1002
ATTR_LAST_KERNEL = 0x1003
ATTR_LAST_DEVICE_US = 0x1004
ATTR_STREAM = 0x1005
ATTR_TIMING = 0x1006
ATTR_AUTOTUNE = 0x1007
ATTR_SELF_CHECK = 0x1008
ATTR_LAST_REUSED = 0x1009
ATTR_LAST_KERNEL_NAME = 0x100A
ATTR_TILES = 0x100B
# this library's error codes (tapp_amd.h)
AMD_ERR_SCRATCH = 108
AMD_ERR_UNSUPPORTED, AMD_ERR_DEVICE, AMD_ERR_NOMEM, AMD_ERR_INVALID, AMD_ERR_KEY, AMD_ERR_BUDGET, AMD_ERR_BUSY = \
    range(101, 108)
KERNEL_AUTO, KERNEL_SIMPLE, KERNEL_MFMA, KERNEL_SPLITK, KERNEL_GROUP, KERNEL_MFMA_SMALL = 0, 1, 2, 3, 4, 5

TYPE_SIZE = {F32: 4, F64: 8, C32: 8, C64: 16, F16: 2, BF16: 2}

_i64p = ctypes.POINTER(ctypes.c_int64)
_H = ctypes.c_ssize_t  # intptr_t


def _single_hip_runtime():
    """PyTorch's ROCm wheel carries its own libamdhip64.so.  If this library
    were loaded first it would bring /opt/rocm's runtime, and a later
    `import torch` would add a second HIP runtime to the process, whose
    contexts do not see each other's allocations.  Importing torch first makes
    the dynamic loader resolve this library's libamdhip64.so.7 to the copy
    already in memory, so one runtime serves both."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def _arr64(vals):
    vals = [int(v) for v in vals]
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)


class TappLib:
    """A loaded TAPP implementation.  `path` defaults to this package's GPU
    library; loading is RTLD_LOCAL so two implementations can share a
    process."""

    API = (
        "TAPP_create_tensor_info", "TAPP_destroy_tensor_info", "TAPP_get_nmodes", "TAPP_set_nmodes",
        "TAPP_get_extents", "TAPP_set_extents", "TAPP_get_strides", "TAPP_set_strides",
        "TAPP_create_tensor_product", "TAPP_destroy_tensor_product", "TAPP_execute_product",
        "TAPP_create_executor", "TAPP_destroy_executor", "TAPP_create_handle", "TAPP_destroy_handle",
        "TAPP_check_success", "TAPP_explain_error",
    )
    EXTENSIONS = (
        "TAPP_execute_batched_product", "TAPP_destroy_status", "TAPP_attr_set", "TAPP_attr_get",
        "TAPP_attr_clear", "TAPP_AMD_version", "TAPP_AMD_plan_info", "TAPP_AMD_last_error_message",
        "TAPP_AMD_create_product_group", "TAPP_AMD_execute_product_group", "TAPP_AMD_destroy_product_group",
        "TAPP_AMD_group_info", "TAPP_AMD_create_server", "TAPP_AMD_server_submit", "TAPP_AMD_status_wait",
        "TAPP_AMD_server_cancel", "TAPP_AMD_server_set_paused", "TAPP_AMD_server_drain", "TAPP_AMD_server_stats",
        "TAPP_AMD_destroy_server", "TAPP_AMD_make_resident", "TAPP_AMD_update_resident",
        "TAPP_AMD_release_resident", "TAPP_AMD_resident_info", "TAPP_AMD_execute_product_noalloc",
    )

    def __init__(self, path=None):
        self.path = path or LIB_PATH
        if not os.path.exists(self.path):
            raise FileNotFoundError(
                f"{self.path} is not built: run `make -C reference-implementation_amd` or __graft_entry__.build()")
        if self.path == LIB_PATH:
            _single_hip_runtime()
        self.lib = ctypes.CDLL(self.path, mode=os.RTLD_LOCAL | os.RTLD_NOW)
        L = self.lib
        L.TAPP_create_tensor_info.argtypes = [ctypes.POINTER(_H), ctypes.c_int, ctypes.c_int, _i64p, _i64p]
        L.TAPP_destroy_tensor_info.argtypes = [_H]
        L.TAPP_get_nmodes.argtypes = [_H]
        L.TAPP_set_nmodes.argtypes = [_H, ctypes.c_int]
        L.TAPP_get_extents.argtypes = [_H, _i64p]
        L.TAPP_get_extents.restype = None
        L.TAPP_set_extents.argtypes = [_H, _i64p]
        L.TAPP_get_strides.argtypes = [_H, _i64p]
        L.TAPP_get_strides.restype = None
        L.TAPP_set_strides.argtypes = [_H, _i64p]
        L.TAPP_create_tensor_product.argtypes = [ctypes.POINTER(_H), _H,
                                                 ctypes.c_int, _H, _i64p, ctypes.c_int, _H, _i64p,
                                                 ctypes.c_int, _H, _i64p, ctypes.c_int, _H, _i64p, ctypes.c_int]
        L.TAPP_destroy_tensor_product.argtypes = [_H]
        L.TAPP_execute_product.argtypes = [_H, _H, ctypes.c_void_p] + [ctypes.c_void_p] * 6
        L.TAPP_create_executor.argtypes = [ctypes.POINTER(_H)]
        L.TAPP_destroy_executor.argtypes = [_H]
        L.TAPP_create_handle.argtypes = [ctypes.POINTER(_H)]
        L.TAPP_destroy_handle.argtypes = [_H]
        L.TAPP_check_success.argtypes = [ctypes.c_int]
        L.TAPP_check_success.restype = ctypes.c_bool
        L.TAPP_explain_error.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_char_p]
        L.TAPP_explain_error.restype = ctypes.c_size_t
        self.has_extensions = hasattr(L, "TAPP_AMD_version")
        if self.has_extensions:
            L.TAPP_execute_batched_product.argtypes = [_H, _H, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                       ctypes.POINTER(ctypes.c_void_p),
                                                       ctypes.POINTER(ctypes.c_void_p), ctypes.c_void_p,
                                                       ctypes.POINTER(ctypes.c_void_p),
                                                       ctypes.POINTER(ctypes.c_void_p)]
            L.TAPP_destroy_status.argtypes = [_H]
            L.TAPP_attr_set.argtypes = [_H, ctypes.c_int, ctypes.c_void_p]
            L.TAPP_attr_get.argtypes = [_H, ctypes.c_int, ctypes.c_void_p]
            L.TAPP_attr_clear.argtypes = [_H, ctypes.c_int]
            L.TAPP_AMD_version.restype = ctypes.c_char_p
            L.TAPP_AMD_last_error_message.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
            L.TAPP_AMD_last_error_message.restype = ctypes.c_size_t
            L.TAPP_AMD_plan_info.argtypes = [_H, _H, ctypes.c_void_p, ctypes.c_void_p, _i64p,
                                             ctypes.POINTER(ctypes.c_int)]
            VP = ctypes.POINTER(ctypes.c_void_p)
            L.TAPP_AMD_create_product_group.argtypes = [ctypes.POINTER(_H), _H, ctypes.c_int, ctypes.POINTER(_H),
                                                        VP, VP, VP, VP, VP, VP]
            L.TAPP_AMD_execute_product_group.argtypes = [_H]
            L.TAPP_AMD_destroy_product_group.argtypes = [_H]
            L.TAPP_AMD_group_info.argtypes = [_H, _i64p]
            L.TAPP_AMD_create_server.argtypes = [ctypes.POINTER(_H), _H, ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int]
            L.TAPP_AMD_server_submit.argtypes = [_H, _H] + [ctypes.c_void_p] * 6 + [ctypes.POINTER(_H)]
            L.TAPP_AMD_status_wait.argtypes = [_H, _H, ctypes.POINTER(ctypes.c_int)]
            L.TAPP_AMD_server_cancel.argtypes = [_H, _H]
            L.TAPP_AMD_server_set_paused.argtypes = [_H, ctypes.c_int]
            L.TAPP_AMD_server_drain.argtypes = [_H]
            L.TAPP_AMD_server_stats.argtypes = [_H, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
            L.TAPP_AMD_destroy_server.argtypes = [_H]
            L.TAPP_AMD_make_resident.argtypes = [_H, ctypes.c_void_p, ctypes.c_size_t]
            L.TAPP_AMD_update_resident.argtypes = [_H, ctypes.c_void_p]
            L.TAPP_AMD_release_resident.argtypes = [_H, ctypes.c_void_p]
            L.TAPP_AMD_resident_info.argtypes = [_H, _i64p]
            L.TAPP_AMD_execute_product_noalloc.argtypes = [_H, _H, ctypes.c_void_p] + [ctypes.c_void_p] * 6

    # --- objects ---------------------------------------------------------
    def create_info(self, dtype, extents, strides):
        h = _H()
        rc = self.lib.TAPP_create_tensor_info(ctypes.byref(h), int(dtype), len(extents), _arr64(extents),
                                              _arr64(strides))
        return rc, h.value

    def destroy_info(self, h):
        return self.lib.TAPP_destroy_tensor_info(h)

    def create_plan(self, ops, infos, idxs, prec, handle=0):
        h = _H()
        args = []
        for op, info, idx in zip(ops, infos, idxs):
            args += [int(op), info, _arr64(idx)]
        rc = self.lib.TAPP_create_tensor_product(ctypes.byref(h), handle, *args, int(prec))
        return rc, h.value

    def destroy_plan(self, h):
        return self.lib.TAPP_destroy_tensor_product(h)

    def create_executor(self):
        h = _H()
        rc = self.lib.TAPP_create_executor(ctypes.byref(h))
        return rc, h.value

    def destroy_executor(self, h):
        return self.lib.TAPP_destroy_executor(h)

    def create_handle(self):
        h = _H()
        rc = self.lib.TAPP_create_handle(ctypes.byref(h))
        return rc, h.value

    def destroy_handle(self, h):
        return self.lib.TAPP_destroy_handle(h)

    def execute(self, plan, exec_, alpha_ptr, a_ptr, b_ptr, beta_ptr, c_ptr, d_ptr, status_ptr=None):
        return self.lib.TAPP_execute_product(plan, exec_, status_ptr, alpha_ptr, a_ptr, b_ptr, beta_ptr,
                                             c_ptr, d_ptr)

    def execute_noalloc(self, plan, exec_, alpha_ptr, a_ptr, b_ptr, beta_ptr, c_ptr, d_ptr):
        """TAPP_AMD_execute_product_noalloc: TAPP_execute_product on the
        device scratch array (no device allocation)."""
        return self.lib.TAPP_AMD_execute_product_noalloc(plan, exec_, None, alpha_ptr, a_ptr, b_ptr, beta_ptr,
                                                         c_ptr, d_ptr)

    def last_error_message(self):
        buf = ctypes.create_string_buffer(512)
        self.lib.TAPP_AMD_last_error_message(buf, 512)
        return buf.value.decode()

    @staticmethod
    def pointer_array(ptrs):
        """A C array of pointers (reusable across calls)."""
        return (ctypes.c_void_p * max(1, len(ptrs)))(*ptrs)

    def execute_batched(self, plan, exec_, alpha_ptr, a_ptrs, b_ptrs, beta_ptr, c_ptrs, d_ptrs):
        """Pointer lists or arrays made by pointer_array."""
        n = len(a_ptrs)
        arr = lambda ps: ps if isinstance(ps, ctypes.Array) else (ctypes.c_void_p * max(1, n))(*ps)  # noqa: E731
        c_arr = arr(c_ptrs) if c_ptrs is not None else None
        return self.lib.TAPP_execute_batched_product(plan, exec_, None, n, alpha_ptr, arr(a_ptrs), arr(b_ptrs),
                                                     beta_ptr, c_arr, arr(d_ptrs))

    def explain(self, rc):
        n = self.lib.TAPP_explain_error(rc, 0, None)
        buf = ctypes.create_string_buffer(n + 1)
        self.lib.TAPP_explain_error(rc, n + 1, buf)
        text = buf.value.decode()
        if self.has_extensions and rc == 102:
            m = ctypes.create_string_buffer(300)
            self.lib.TAPP_AMD_last_error_message(m, 300)
            text += " " + m.value.decode()
        return text

    # --- extensions ---------------------------------------------------------
    def attr_set_int(self, exec_, key, value):
        v = ctypes.c_int(int(value))
        return self.lib.TAPP_attr_set(exec_, key, ctypes.byref(v))

    def attr_get_int(self, exec_, key):
        v = ctypes.c_int(0)
        rc = self.lib.TAPP_attr_get(exec_, key, ctypes.byref(v))
        return rc, v.value

    def attr_get_double(self, exec_, key):
        v = ctypes.c_double(0)
        rc = self.lib.TAPP_attr_get(exec_, key, ctypes.byref(v))
        return rc, v.value

    def last_kernel_name(self, exec_):
        v = ctypes.c_char_p()
        rc = self.lib.TAPP_attr_get(exec_, ATTR_LAST_KERNEL_NAME, ctypes.byref(v))
        return rc, (v.value or b"").decode()

    def create_group(self, exec_, plans, alphas, a_ptrs, b_ptrs, betas, c_ptrs, d_ptrs):
        n = len(plans)
        arr = lambda ps: (ctypes.c_void_p * max(1, n))(*ps)  # noqa: E731
        h = _H()
        rc = self.lib.TAPP_AMD_create_product_group(ctypes.byref(h), exec_, n, (_H * max(1, n))(*plans),
                                                    arr(alphas), arr(a_ptrs), arr(b_ptrs), arr(betas),
                                                    arr(c_ptrs), arr(d_ptrs))
        return rc, h.value

    def execute_group(self, g):
        return self.lib.TAPP_AMD_execute_product_group(g)

    def destroy_group(self, g):
        return self.lib.TAPP_AMD_destroy_product_group(g)

    def group_info(self, g):
        info = (ctypes.c_int64 * 3)()
        rc = self.lib.TAPP_AMD_group_info(g, info)
        return rc, list(info)

    # --- serving -------------------------------------------------------------
    SERVER_PAUSED = 1
    STAT_NAMES = ("completed", "p50_us", "p99_us", "mean_us", "stalls", "max_inflight", "batches",
                  "persistent_products", "gflops", "queued", "held_bytes", "max_held_bytes", "reclaims")

    def create_server(self, exec_, budget_bytes, slots=4, max_batch=64, flags=0):
        h = _H()
        rc = self.lib.TAPP_AMD_create_server(ctypes.byref(h), exec_, int(budget_bytes), slots, max_batch, flags)
        return rc, h.value

    def submit(self, srv, plan, alpha_ptr, a_ptr, b_ptr, beta_ptr, c_ptr, d_ptr):
        st = _H()
        rc = self.lib.TAPP_AMD_server_submit(srv, plan, alpha_ptr, a_ptr, b_ptr, beta_ptr, c_ptr, d_ptr,
                                             ctypes.byref(st))
        return rc, st.value

    def wait(self, srv, status):
        r = ctypes.c_int(-1)
        rc = self.lib.TAPP_AMD_status_wait(srv, status, ctypes.byref(r))
        return rc, r.value

    def cancel(self, srv, status):
        return self.lib.TAPP_AMD_server_cancel(srv, status)

    def destroy_status(self, status):
        return self.lib.TAPP_destroy_status(status)

    def set_paused(self, srv, paused):
        return self.lib.TAPP_AMD_server_set_paused(srv, int(paused))

    def drain(self, srv):
        return self.lib.TAPP_AMD_server_drain(srv)

    def server_stats(self, srv, reset=False):
        v = (ctypes.c_double * len(self.STAT_NAMES))()
        rc = self.lib.TAPP_AMD_server_stats(srv, v, int(reset))
        return rc, dict(zip(self.STAT_NAMES, list(v)))

    def destroy_server(self, srv):
        return self.lib.TAPP_AMD_destroy_server(srv)

    # resident tensors
    def make_resident(self, handle, host_ptr, nbytes):
        return self.lib.TAPP_AMD_make_resident(handle, host_ptr, nbytes)

    def update_resident(self, handle, host_ptr):
        return self.lib.TAPP_AMD_update_resident(handle, host_ptr)

    def release_resident(self, handle, host_ptr):
        return self.lib.TAPP_AMD_release_resident(handle, host_ptr)

    def resident_info(self, handle):
        info = (ctypes.c_int64 * 2)()
        rc = self.lib.TAPP_AMD_resident_info(handle, info)
        return rc, list(info)

    def plan_info(self, plan, exec_, beta_ptr, c_ptr):
        sizes = (ctypes.c_int64 * 6)()
        ct = ctypes.c_int(-1)
        rc = self.lib.TAPP_AMD_plan_info(plan, exec_, beta_ptr, c_ptr, sizes, ctypes.byref(ct))
        return rc, list(sizes), ct.value


def trace_replay(lib, srv, plans, alphas, a_ptrs, b_ptrs, betas, c_ptrs, d_ptrs, arrival_us):
    """Open-loop submit of request i at arrival_us[i] (microseconds after the
    start) through the serving loop, waiting for all.  Returns (failed
    submits, per-request codes, total submit lateness in microseconds)."""
    _single_hip_runtime()
    tr = ctypes.CDLL(TRACE_PATH, mode=os.RTLD_LOCAL)
    n = len(plans)
    VP = ctypes.c_void_p * max(1, n)
    tr.tapp_trace_replay.argtypes = [_H, ctypes.c_int, ctypes.POINTER(_H)] + [ctypes.POINTER(ctypes.c_void_p)] * 6 + [
        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    rcs = (ctypes.c_int * max(1, n))()
    late = ctypes.c_double(0)
    arr = (ctypes.c_double * max(1, n))(*arrival_us)
    failed = tr.tapp_trace_replay(srv, n, (_H * max(1, n))(*plans), VP(*alphas), VP(*a_ptrs), VP(*b_ptrs),
                                  VP(*betas), VP(*c_ptrs), VP(*d_ptrs), arr, rcs, ctypes.byref(late))
    return failed, list(rcs)[:n], late.value


def trace_sequential(lib, exec_, plans, alphas, a_ptrs, b_ptrs, betas, c_ptrs, d_ptrs):
    """The same requests as one TAPP_execute_product call each, back to back
    on one executor (from C).  Returns (failed calls, per-request codes,
    elapsed seconds)."""
    _single_hip_runtime()
    tr = ctypes.CDLL(TRACE_PATH, mode=os.RTLD_LOCAL)
    n = len(plans)
    VP = ctypes.c_void_p * max(1, n)
    tr.tapp_trace_sequential.argtypes = [_H, ctypes.c_int, ctypes.POINTER(_H)] + [ctypes.POINTER(ctypes.c_void_p)] * 6 + [
        ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_double)]
    rcs = (ctypes.c_int * max(1, n))()
    el = ctypes.c_double(0)
    failed = tr.tapp_trace_sequential(exec_, n, (_H * max(1, n))(*plans), VP(*alphas), VP(*a_ptrs), VP(*b_ptrs),
                                      VP(*betas), VP(*c_ptrs), VP(*d_ptrs), rcs, ctypes.byref(el))
    return failed, list(rcs)[:n], el.value * 1e-6
