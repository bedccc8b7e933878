"""ctypes binding of the gfx950 library (lib/libpt_hip.so, C ABI in
include/pandemic/pt_abi.h).

PandemicSim mirrors the reference's host flow (PandemicThrust/PandemicThrust/
main.cu:12-31): construct, setup_sim(), run_to_completion().  Every call goes
to the HIP library; there is no CPU fallback, and importing raises when the
library is not built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libpt_hip.so")

PT_STATUS_P, PT_STATUS_S, PT_DAY_P, PT_DAY_S, PT_GEN_P, PT_GEN_S = range(6)
PT_AGE, PT_HOUSEHOLD, PT_HH_OFFSETS, PT_WP_OFFSETS, PT_WP_PEOPLE = range(6, 11)
PT_INFECTED, PT_ERRAND_OFFSETS, PT_ERRAND_PEOPLE, PT_HH_TYPES, PT_WORKPLACE, PT_ERRANDS = range(11, 17)
PT_LOCATION_TYPE_OFFSETS = 17
_U8_ARRAYS = {PT_STATUS_P, PT_STATUS_S, PT_DAY_P, PT_DAY_S, PT_GEN_P, PT_GEN_S, PT_AGE, PT_HH_TYPES}

EXPORTS = [
    "pt_last_error", "pt_glibc_seeds", "pt_glibc_initial_infected", "pt_sim_create", "pt_sim_setup",
    "pt_sim_seed_infections", "pt_sim_step", "pt_sim_run", "pt_sim_generation_offsets", "pt_sim_array_bytes",
    "pt_sim_copy_array", "pt_sim_get_scalars", "pt_sim_get_layout", "pt_sim_set_table_mode", "pt_sim_table_log",
    "pt_sim_snapshot",
    "pt_sim_restore", "pt_sim_enable_timing", "pt_sim_set_interventions", "pt_batch_set_interventions",
    "pt_sim_kernel_times", "pt_sim_kernel_name", "pt_sim_synchronize", "pt_sim_last_run_ms", "pt_sim_destroy",
    "pt_batch_create", "pt_batch_setup", "pt_batch_seed_infections", "pt_batch_get_info", "pt_batch_run",
    "pt_batch_generation_offsets", "pt_batch_array_bytes", "pt_batch_copy_array", "pt_batch_set_table_mode",
    "pt_batch_snapshot", "pt_batch_restore", "pt_batch_synchronize", "pt_batch_last_run_ms", "pt_batch_destroy",
    "pt_mlp_param_count", "pt_mlp_create", "pt_mlp_set_params", "pt_mlp_get_params", "pt_mlp_set_data",
    "pt_mlp_train", "pt_mlp_gradients", "pt_mlp_predict", "pt_mlp_last_train_ms", "pt_mlp_destroy",
    "pt_sim_day_kernel_times", "pt_sim_setup_times", "pt_sim_daily_rows", "pt_sim_checkpoint_bytes",
    "pt_sim_checkpoint_save", "pt_sim_checkpoint_load", "pt_device_memory", "pt_device_name",
    "pt_batch_use_graphs", "pt_sim_set_launch", "pt_sim_get_launch", "pt_batch_set_launch", "pt_batch_get_launch",
    "pt_sim_enable_trace", "pt_sim_trace_size", "pt_sim_trace", "pt_sim_trace_draws", "pt_location_table",
    "pt_location_table_noalloc", "pt_mlp_use_graphs",
]

# include/pandemic/pt_abi.h: largest input of pt_location_table_noalloc
PT_LOCATION_TABLE_NOALLOC_MAX_PAIRS = 4194304

# enum pt_launch_knob (include/pandemic/pt_abi.h)
LAUNCH_KNOBS = ["contact_blocks_per_cu", "gen_blocks_per_cu", "sort_wide_bits", "mark_limit_pct", "subset_share_pct",
                "sort_max_bits", "segmented_tables"]

# enum pt_setup_phase (include/pandemic/pt_abi.h)
SETUP_PHASES = ["scale", "push", "population_size", "size_arrays", "households", "workplaces", "status",
                "initial_infected"]


class Params(C.Structure):
    """The data row of constants.csv (PandemicSim.cu:299-309)."""
    _fields_ = [
        ("max_days", C.c_int32),
        ("r_pandemic", C.c_float),
        ("r_seasonal", C.c_float),
        ("init_pandemic", C.c_int32),
        ("init_seasonal", C.c_int32),
        ("people_scaling", C.c_float),
        ("location_scaling", C.c_float),
        ("percent_symptomatic", C.c_float),
        ("asymp_factor", C.c_float),
    ]

    @classmethod
    def from_row(cls, fields):
        f = [x.strip() for x in fields]
        return cls(int(f[0]), float(f[1]), float(f[2]), int(f[3]), int(f[4]), float(f[5]), float(f[6]),
                   float(f[7]), float(f[8]))


class Interventions(C.Structure):
    """pt_interventions (include/pandemic/pt_abi.h); the defaults are the
    neutral set, which reproduces the reference exactly."""
    _fields_ = [
        ("latent_days", C.c_int32 * 2),
        ("infectious_days", C.c_int32 * 2),
        ("school_close_start", C.c_int32),
        ("school_close_days", C.c_int32),
        ("workplace_reduction", C.c_float),
        ("errand_reduction", C.c_float),
        ("vaccination", C.c_float * 2),
    ]

    def __init__(self, latent=(0, 0), infectious=(10, 10), school_close_start=0, school_close_days=0,
                 workplace_reduction=0.0, errand_reduction=0.0, vaccination=(0.0, 0.0)):
        super().__init__((C.c_int32 * 2)(*latent), (C.c_int32 * 2)(*infectious), school_close_start,
                         school_close_days, workplace_reduction, errand_reduction, (C.c_float * 2)(*vaccination))


class PTError(RuntimeError):
    pass


_lib = None


def lib():
    """Loads the HIP library (raises ImportError if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpt_hip.so is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        L.pt_last_error.restype = C.c_char_p
        L.pt_glibc_seeds.argtypes = [C.c_int32, C.c_void_p]
        L.pt_glibc_initial_infected.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.pt_sim_create.argtypes = [C.POINTER(Params), C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.pt_sim_setup.argtypes = [C.c_void_p, C.POINTER(C.c_int64)]
        L.pt_sim_seed_infections.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32]
        L.pt_sim_step.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
        L.pt_sim_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]
        L.pt_sim_generation_offsets.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_array_bytes.restype = C.c_int64
        L.pt_sim_array_bytes.argtypes = [C.c_void_p, C.c_int32]
        L.pt_sim_copy_array.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64]
        L.pt_sim_get_scalars.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_sim_get_layout.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_sim_set_table_mode.argtypes = [C.c_void_p, C.c_int32]
        L.pt_sim_set_interventions.argtypes = [C.c_void_p, C.POINTER(Interventions)]
        L.pt_batch_set_interventions.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_sim_table_log.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_snapshot.argtypes = [C.c_void_p]
        L.pt_sim_restore.argtypes = [C.c_void_p]
        L.pt_sim_enable_timing.argtypes = [C.c_void_p, C.c_int32]
        L.pt_sim_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_kernel_name.restype = C.c_char_p
        L.pt_sim_kernel_name.argtypes = [C.c_int32]
        L.pt_sim_synchronize.argtypes = [C.c_void_p]
        L.pt_sim_last_run_ms.restype = C.c_double
        L.pt_sim_last_run_ms.argtypes = [C.c_void_p]
        L.pt_sim_destroy.argtypes = [C.c_void_p]
        L.pt_batch_create.argtypes = [C.c_int32, C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(C.c_void_p)]
        L.pt_batch_setup.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_batch_seed_infections.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pt_batch_get_info.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_batch_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pt_batch_generation_offsets.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_batch_array_bytes.restype = C.c_int64
        L.pt_batch_array_bytes.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        L.pt_batch_copy_array.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64]
        L.pt_batch_set_table_mode.argtypes = [C.c_void_p, C.c_int32]
        L.pt_batch_snapshot.argtypes = [C.c_void_p]
        L.pt_batch_restore.argtypes = [C.c_void_p]
        L.pt_batch_synchronize.argtypes = [C.c_void_p]
        L.pt_batch_last_run_ms.restype = C.c_double
        L.pt_batch_last_run_ms.argtypes = [C.c_void_p]
        L.pt_batch_destroy.argtypes = [C.c_void_p]
        L.pt_mlp_param_count.restype = C.c_int64
        L.pt_mlp_create.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        L.pt_mlp_set_params.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_mlp_get_params.argtypes = [C.c_void_p, C.c_void_p]
        L.pt_mlp_set_data.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.pt_mlp_train.argtypes = [C.c_void_p, C.c_void_p, C.c_int32, C.c_float, C.c_void_p]
        L.pt_mlp_gradients.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.pt_mlp_predict.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]
        L.pt_mlp_last_train_ms.restype = C.c_double
        L.pt_mlp_last_train_ms.argtypes = [C.c_void_p]
        L.pt_mlp_use_graphs.argtypes = [C.c_void_p, C.c_int32]
        L.pt_mlp_destroy.argtypes = [C.c_void_p]
        L.pt_sim_day_kernel_times.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_setup_times.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_daily_rows.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.pt_sim_checkpoint_bytes.restype = C.c_int64
        L.pt_sim_checkpoint_bytes.argtypes = [C.c_void_p]
        L.pt_sim_checkpoint_save.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.pt_sim_checkpoint_load.argtypes = [C.c_void_p, C.c_void_p, C.c_int64]
        L.pt_device_memory.argtypes = [C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.pt_device_name.argtypes = [C.c_int32, C.c_char_p, C.c_int32]
        L.pt_batch_use_graphs.argtypes = [C.c_void_p, C.c_int32]
        L.pt_sim_enable_trace.argtypes = [C.c_void_p, C.c_int32]
        L.pt_sim_trace_size.restype = C.c_int64
        L.pt_sim_trace_size.argtypes = [C.c_void_p]
        L.pt_sim_trace.argtypes = [C.c_void_p] + [C.c_void_p] * 5 + [C.c_int64]
        L.pt_sim_trace_draws.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64]
        L.pt_location_table.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_void_p]
        L.pt_location_table_noalloc.argtypes = L.pt_location_table.argtypes
        for f in ("pt_sim_set_launch", "pt_batch_set_launch"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int32, C.c_int32]
        for f in ("pt_sim_get_launch", "pt_batch_get_launch"):
            getattr(L, f).argtypes = [C.c_void_p, C.c_int32]
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise PTError(lib().pt_last_error().decode())
    return rc


def glibc_seeds(core_seed):
    out = (C.c_int32 * 4)()
    _check(lib().pt_glibc_seeds(core_seed, out))
    return list(out)


def device_memory(device=0):
    """(free bytes, total bytes) of a device."""
    f, t = C.c_uint64(0), C.c_uint64(0)
    _check(lib().pt_device_memory(device, C.byref(f), C.byref(t)))
    return int(f.value), int(t.value)


def device_name(device=0):
    buf = C.create_string_buffer(256)
    _check(lib().pt_device_name(device, buf, 256))
    return buf.value.decode(errors="replace")


def glibc_initial_infected(core_seed, n_people, n):
    out = np.zeros(max(n, 1), np.int32)
    _check(lib().pt_glibc_initial_infected(core_seed, n_people, n, out.ctypes.data))
    return out[:n]


class PandemicSim:
    """One simulation on one GPU (the reference's PandemicSim object)."""

    def __init__(self, params, core_seed, device=0):
        self.params = params
        self.core_seed = core_seed
        self.seeds = glibc_seeds(core_seed)
        key = (C.c_int32 * 2)(self.seeds[0], self.seeds[1])
        h = C.c_void_p()
        _check(lib().pt_sim_create(C.byref(params), key, device, C.byref(h)))
        self.h = h
        self.n_people = 0

    def close(self):
        if getattr(self, "h", None):
            lib().pt_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # setupSim (PandemicSim.cu:146-208)
    def setup_sim(self):
        n = C.c_int64(0)
        _check(lib().pt_sim_setup(self.h, C.byref(n)))
        self.n_people = int(n.value)
        p = self.params
        ids = glibc_initial_infected(self.core_seed, self.n_people, p.init_pandemic + p.init_seasonal)
        _check(lib().pt_sim_seed_infections(self.h, ids.ctypes.data, p.init_pandemic, p.init_seasonal))
        return self.n_people

    def step(self):
        counts = np.zeros(16, np.int32)
        inf = C.c_int32(0)
        rc = _check(lib().pt_sim_step(self.h, counts.ctypes.data, C.byref(inf)))
        return rc, counts, int(inf.value)

    # runToCompletion loop (PandemicSim.cu:877-967)
    def run(self):
        md = max(self.params.max_days, 1)
        counts = np.zeros((md, 16), np.int32)
        inf = np.zeros(md, np.int32)
        days = C.c_int32(0)
        _check(lib().pt_sim_run(self.h, counts.ctypes.data, inf.ctypes.data, C.byref(days)))
        d = int(days.value)
        return counts[:d].copy(), inf[:d].copy()

    def setup_only(self):
        """setupSim without seeding, for load_checkpoint()."""
        n = C.c_int64(0)
        _check(lib().pt_sim_setup(self.h, C.byref(n)))
        self.n_people = int(n.value)
        return self.n_people

    def daily_rows(self):
        """Every daily row since setup (restored rows included)."""
        counts = np.zeros((256, 16), np.int32)
        inf = np.zeros(256, np.int32)
        n = _check(lib().pt_sim_daily_rows(self.h, counts.ctypes.data, inf.ctypes.data, 256))
        return counts[:n].copy(), inf[:n].copy()

    def save_checkpoint(self):
        """The loop state at the current day boundary, as bytes."""
        n = int(lib().pt_sim_checkpoint_bytes(self.h))
        if n < 0:
            raise PTError("checkpoint before setup")
        buf = np.zeros(n, np.uint8)
        _check(lib().pt_sim_checkpoint_save(self.h, buf.ctypes.data, n))
        return buf.tobytes()

    def load_checkpoint(self, data):
        buf = np.frombuffer(data, np.uint8)
        _check(lib().pt_sim_checkpoint_load(self.h, buf.ctypes.data, len(buf)))

    def enable_trace(self, on=True):
        """Per-contact validation trace of every contact day (pt_sim_trace)."""
        _check(lib().pt_sim_enable_trace(self.h, 1 if on else 0))

    def trace(self):
        """The last contact day's trace: infector, victim, type, loc, action per slot."""
        n = int(lib().pt_sim_trace_size(self.h))
        if n < 0:
            raise PTError(lib().pt_last_error().decode())
        out = {"infector": np.zeros(n, np.int32), "victim": np.zeros(n, np.int32), "type": np.zeros(n, np.uint8),
               "loc": np.zeros(n, np.int32), "action": np.zeros(n, np.uint8)}
        if n:
            _check(lib().pt_sim_trace(self.h, out["infector"].ctypes.data, out["victim"].ctypes.data,
                                      out["type"].ctypes.data, out["loc"].ctypes.data, out["action"].ctypes.data, n))
        return out

    def trace_draws(self):
        """The last contact day's draws per slot (pt_sim_trace_draws): draw bits
        (2 pandemic, 1 seasonal succeeded) and profile draws (low nibble
        pandemic, high nibble seasonal)."""
        n = int(lib().pt_sim_trace_size(self.h))
        if n < 0:
            raise PTError(lib().pt_last_error().decode())
        draw = np.zeros(n, np.uint8)
        prof = np.zeros(n, np.uint8)
        if n:
            _check(lib().pt_sim_trace_draws(self.h, draw.ctypes.data, prof.ctypes.data, n))
        return draw, prof

    def set_launch(self, name, value):
        """Launch shape (LAUNCH_KNOBS); speed only, results unchanged."""
        _check(lib().pt_sim_set_launch(self.h, LAUNCH_KNOBS.index(name), int(value)))

    def launch(self):
        return {k: _check(lib().pt_sim_get_launch(self.h, i)) for i, k in enumerate(LAUNCH_KNOBS)}

    def setup_times(self):
        ms = np.zeros(len(SETUP_PHASES), np.float64)
        _check(lib().pt_sim_setup_times(self.h, ms.ctypes.data, len(ms)))
        return dict(zip(SETUP_PHASES, (float(x) for x in ms)))

    def day_kernel_times(self):
        """days x classes milliseconds since timing was enabled."""
        k = len(self.kernel_classes())
        days = _check(lib().pt_sim_day_kernel_times(self.h, None, 0))
        ms = np.zeros((max(days, 1), k), np.float64)
        _check(lib().pt_sim_day_kernel_times(self.h, ms.ctypes.data, days))
        return ms[:days]

    def generation_offsets(self):
        md = self.params.max_days
        cp = np.zeros(md + 1, np.int32)
        cs = np.zeros(md + 1, np.int32)
        _check(lib().pt_sim_generation_offsets(self.h, cp.ctypes.data, cs.ctypes.data, md + 1))
        return cp, cs

    def array(self, which):
        nbytes = int(lib().pt_sim_array_bytes(self.h, which))
        if nbytes < 0:
            raise PTError(lib().pt_last_error().decode())
        dt = np.uint8 if which in _U8_ARRAYS else np.int32
        out = np.zeros(nbytes // np.dtype(dt).itemsize, dt)
        if nbytes:
            _check(lib().pt_sim_copy_array(self.h, which, out.ctypes.data, nbytes))
        return out

    def scalars(self):
        out = np.zeros(8, np.int64)
        _check(lib().pt_sim_get_scalars(self.h, out.ctypes.data))
        return out

    def layout(self):
        out = np.zeros(8, np.int64)
        _check(lib().pt_sim_get_layout(self.h, out.ctypes.data))
        keys = ["wd_base", "wd_E", "we_base", "we_E", "wd_buckets", "we_buckets", "wd_passes", "we_passes"]
        return dict(zip(keys, (int(x) for x in out)))

    def set_table_mode(self, mode):
        """0 automatic, 1 full errand tables every day, 2 queried buckets only."""
        _check(lib().pt_sim_set_table_mode(self.h, mode))

    def set_interventions(self, iv):
        """Interventions (an Interventions); call before setup_sim()."""
        _check(lib().pt_sim_set_interventions(self.h, C.byref(iv)))

    def table_log(self):
        """Per day since setup / restore: (full table?, pairs sorted)."""
        full = np.zeros(256, np.int32)
        pairs = np.zeros(256, np.int32)
        n = _check(lib().pt_sim_table_log(self.h, full.ctypes.data, pairs.ctypes.data, 256))
        return full[:n].astype(bool), pairs[:n].astype(np.int64)

    def snapshot(self):
        _check(lib().pt_sim_snapshot(self.h))

    def restore(self):
        _check(lib().pt_sim_restore(self.h))

    def enable_timing(self, on=True, only=None):
        """Times every kernel class, or only the class named `only`."""
        if not on:
            mode = 0
        elif only is None:
            mode = -1
        else:
            mode = self.kernel_classes().index(only) + 1
        _check(lib().pt_sim_enable_timing(self.h, mode))

    def kernel_classes(self):
        n = _check(lib().pt_sim_kernel_times(self.h, None, None, 0))
        return [lib().pt_sim_kernel_name(i).decode() for i in range(n)]

    def kernel_times(self):
        n = 32
        ms = np.zeros(n, np.float64)
        cnt = np.zeros(n, np.int64)
        k = _check(lib().pt_sim_kernel_times(self.h, ms.ctypes.data, cnt.ctypes.data, n))
        return {lib().pt_sim_kernel_name(i).decode(): (float(ms[i]), int(cnt[i])) for i in range(k)}

    def synchronize(self):
        _check(lib().pt_sim_synchronize(self.h))

    def last_run_ms(self):
        return float(lib().pt_sim_last_run_ms(self.h))


class PandemicBatch:
    """Many independent simulations (replicates) in lockstep on one GPU
    (include/pandemic/pt_batch.h).  Replicate r ends exactly where a
    PandemicSim with rows[r] and core_seeds[r] ends; the reference runs one
    process per constants row instead (PythonBatcher/write_gpu_constants.py:6-49)."""

    def __init__(self, params_list, core_seeds, device=0):
        n = len(params_list)
        if n != len(core_seeds) or n == 0:
            raise ValueError("one core seed per constants row")
        self.params = list(params_list)
        self.core_seeds = [int(s) for s in core_seeds]
        arr = (Params * n)(*self.params)
        keys = np.zeros(2 * n, np.int32)
        for r, s in enumerate(self.core_seeds):
            sd = glibc_seeds(s)
            keys[2 * r], keys[2 * r + 1] = sd[0], sd[1]
        h = C.c_void_p()
        _check(lib().pt_batch_create(n, arr, keys.ctypes.data, device, C.byref(h)))
        self.h = h
        self.n = n
        self.n_people = None

    def close(self):
        if getattr(self, "h", None):
            lib().pt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def setup(self):
        """setupSim plus setup_initialInfected of every replicate; returns the populations."""
        n_people = np.zeros(self.n, np.int64)
        _check(lib().pt_batch_setup(self.h, n_people.ctypes.data))
        self.n_people = n_people
        ids, n_p, n_s = [], np.zeros(self.n, np.int32), np.zeros(self.n, np.int32)
        for r, p in enumerate(self.params):
            ids.append(glibc_initial_infected(self.core_seeds[r], int(n_people[r]), p.init_pandemic + p.init_seasonal))
            n_p[r], n_s[r] = p.init_pandemic, p.init_seasonal
        flat = np.ascontiguousarray(np.concatenate(ids) if ids else np.zeros(0, np.int32), dtype=np.int32)
        _check(lib().pt_batch_seed_infections(self.h, flat.ctypes.data if flat.size else None, n_p.ctypes.data,
                                              n_s.ctypes.data))
        return n_people

    def info(self):
        out = np.zeros(4, np.int64)
        _check(lib().pt_batch_get_info(self.h, out.ctypes.data))
        return {"replicates": int(out[0]), "D": int(out[1]), "agents": int(out[2]), "slots": int(out[3])}

    def run(self):
        """Runs every replicate to its end: lists of (counts[days, 16], infected[days])."""
        D = self.info()["D"]
        counts = np.zeros((self.n, D, 16), np.int32)
        inf = np.zeros((self.n, D), np.int32)
        days = np.zeros(self.n, np.int32)
        _check(lib().pt_batch_run(self.h, counts.ctypes.data, inf.ctypes.data, days.ctypes.data))
        return [(counts[r, :days[r]].copy(), inf[r, :days[r]].copy()) for r in range(self.n)]

    def generation_offsets(self, r):
        md = self.params[r].max_days
        cp = np.zeros(md + 1, np.int32)
        cs = np.zeros(md + 1, np.int32)
        _check(lib().pt_batch_generation_offsets(self.h, r, cp.ctypes.data, cs.ctypes.data, md + 1))
        return cp, cs

    def array(self, r, which):
        nbytes = int(lib().pt_batch_array_bytes(self.h, r, which))
        if nbytes < 0:
            raise PTError("array %d of replicate %d is not available" % (which, r))
        dt = np.uint8 if which in _U8_ARRAYS else np.int32
        out = np.zeros(nbytes // np.dtype(dt).itemsize, dt)
        if nbytes:
            _check(lib().pt_batch_copy_array(self.h, r, which, out.ctypes.data, nbytes))
        return out

    def set_table_mode(self, mode):
        _check(lib().pt_batch_set_table_mode(self.h, mode))

    def set_launch(self, name, value):
        """Launch shape (LAUNCH_KNOBS); speed only, results unchanged."""
        _check(lib().pt_batch_set_launch(self.h, LAUNCH_KNOBS.index(name), int(value)))

    def launch(self):
        return {k: _check(lib().pt_batch_get_launch(self.h, i)) for i, k in enumerate(LAUNCH_KNOBS)}

    def use_graphs(self, on=True):
        """Day step replayed from HIP graphs (default) or enqueued kernel by kernel."""
        _check(lib().pt_batch_use_graphs(self.h, 1 if on else 0))

    def set_interventions(self, ivs):
        """One Interventions per replicate; call before setup()."""
        arr = (Interventions * self.n)(*ivs)
        _check(lib().pt_batch_set_interventions(self.h, arr))

    def snapshot(self):
        _check(lib().pt_batch_snapshot(self.h))

    def restore(self):
        _check(lib().pt_batch_restore(self.h))

    def synchronize(self):
        _check(lib().pt_batch_synchronize(self.h))

    def last_run_ms(self):
        return float(lib().pt_batch_last_run_ms(self.h))
