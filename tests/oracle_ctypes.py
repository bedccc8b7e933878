"""ctypes binding of the CPU oracle (oracle/build/liboracle.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "build", "liboracle.so")

OR_STATUS_P, OR_STATUS_S, OR_DAY_P, OR_DAY_S, OR_GEN_P, OR_GEN_S = range(6)
OR_AGE, OR_HOUSEHOLD, OR_HH_OFFSETS, OR_WP_OFFSETS, OR_WP_PEOPLE = range(6, 11)
OR_INFECTED, OR_ERRAND_OFFSETS, OR_ERRAND_PEOPLE, OR_HH_TYPES, OR_WP_KEYS = range(11, 16)

_U8 = {OR_STATUS_P, OR_STATUS_S, OR_DAY_P, OR_DAY_S, OR_GEN_P, OR_GEN_S, OR_AGE, OR_HH_TYPES}


class Params(C.Structure):
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


def params_from_row(row):
    """row: the data fields of a constants.csv line after the label."""
    f = [x.strip() for x in row]
    return Params(int(f[0]), float(f[1]), float(f[2]), int(f[3]), int(f[4]),
                  float(f[5]), float(f[6]), float(f[7]), float(f[8]))


_libs = {}


def plant_path(k):
    """The oracle copy with planted mistake k (OR_PLANT, oracle/pt_oracle.cpp)."""
    return os.path.join(ROOT, "oracle", "build", "liboracle_plant%d.so" % k)


def lib(path=LIB_PATH):
    L = _libs.get(path)
    if L is None:
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run __graft_entry__.build() or make -C oracle")
        L = C.CDLL(path)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(Params), C.c_int32]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_last_error.restype = C.c_char_p
        L.or_setup.argtypes = [C.c_void_p]
        L.or_set_interventions.argtypes = [C.c_void_p, C.c_void_p]
        L.or_step.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.or_get_size.restype = C.c_int64
        L.or_get_size.argtypes = [C.c_void_p, C.c_int]
        L.or_get_array.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int64]
        L.or_get_scalar.restype = C.c_int64
        L.or_get_scalar.argtypes = [C.c_void_p, C.c_int]
        L.or_get_type_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_get_infection_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_generation_offsets.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.or_recalc_workplace.restype = C.c_int32
        L.or_recalc_workplace.argtypes = [C.c_void_p, C.c_int32, C.c_uint8]
        L.or_recalc_weekday.restype = C.c_int32
        L.or_recalc_weekday.argtypes = [C.c_void_p, C.c_int32, C.c_uint8, C.c_void_p]
        L.or_recalc_weekend.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.or_set_offsets.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64]
        L.or_fish_workplace.restype = C.c_int32
        L.or_fish_workplace.argtypes = [C.c_void_p, C.c_uint32]
        L.or_fish_school_and_age.argtypes = [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
        L.or_fish_household_type.restype = C.c_uint8
        L.or_fish_household_type.argtypes = [C.c_uint32]
        L.or_threefry2x64_20.argtypes = [C.c_uint64] * 4 + [C.c_void_p]
        L.or_threefry2x32_20.argtypes = [C.c_uint32] * 4 + [C.c_void_p]
        L.or_glibc_seeds.argtypes = [C.c_int32, C.c_void_p]
        L.or_glibc_initial_infected.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_void_p]
        L.or_enable_trace.argtypes = [C.c_void_p, C.c_int]
        L.or_trace_size.restype = C.c_int64
        L.or_trace_size.argtypes = [C.c_void_p]
        L.or_get_trace.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.or_write_outputs.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p, C.c_void_p, C.c_int32]
        L.or_get_trace_draws.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_infection_profile.restype = C.c_uint8
        L.or_infection_profile.argtypes = [C.c_void_p, C.c_uint32]
        L.or_infection_probability.restype = C.c_float
        L.or_infection_probability.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        L.or_plant.restype = C.c_int
        L.or_set_state.argtypes = [C.c_void_p, C.c_int32, C.c_uint64] + [C.c_void_p] * 6
        _libs[path] = L
    return L


def threefry2x64(c0, c1, k0, k1):
    out = (C.c_uint64 * 2)()
    lib().or_threefry2x64_20(c0, c1, k0, k1, out)
    return int(out[0]), int(out[1])


def threefry2x32(c0, c1, k0, k1):
    out = (C.c_uint32 * 2)()
    lib().or_threefry2x32_20(c0, c1, k0, k1, out)
    return int(out[0]), int(out[1])


def glibc_seeds(core_seed):
    out = (C.c_int32 * 4)()
    lib().or_glibc_seeds(core_seed, out)
    return list(out)


class OracleSim:
    """One oracle simulation (mirrors the reference's PandemicSim object)."""

    def __init__(self, params, core_seed, path=LIB_PATH):
        self.params = params
        self.core_seed = core_seed
        self.L = lib(path)
        self.h = self.L.or_create(C.byref(params), core_seed)
        if not self.h:
            raise RuntimeError(self.L.or_last_error().decode())

    def close(self):
        if self.h:
            self.L.or_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_interventions(self, iv):
        """iv: a pandemicthrust_amd.Interventions (same layout as or_interventions)."""
        self.L.or_set_interventions(self.h, C.byref(iv))

    def setup(self):
        if self.L.or_setup(self.h) != 0:
            raise RuntimeError(self.L.or_last_error().decode())

    def step(self):
        counts = np.zeros(16, np.int32)
        inf = C.c_int32(0)
        rc = self.L.or_step(self.h, counts.ctypes.data, C.byref(inf))
        return rc, counts, inf.value

    def run(self, cap_days=256):
        counts = np.zeros((cap_days, 16), np.int32)
        inf = np.zeros(cap_days, np.int32)
        d = self.L.or_run(self.h, counts.ctypes.data, inf.ctypes.data, cap_days)
        return counts[:d].copy(), inf[:d].copy()

    def scalar(self, which):
        return int(self.L.or_get_scalar(self.h, which))

    @property
    def n_people(self):
        return self.scalar(0)

    def array(self, which):
        size = int(self.L.or_get_size(self.h, which))
        dt = np.uint8 if which in _U8 else np.int32
        out = np.zeros(size // np.dtype(dt).itemsize, dt)
        if size:
            rc = self.L.or_get_array(self.h, which, out.ctypes.data, size)
            assert rc == 0
        return out

    def type_tables(self):
        cnt = np.zeros(14, np.int32)
        off = np.zeros(14, np.int32)
        self.L.or_get_type_tables(self.h, cnt.ctypes.data, off.ctypes.data)
        return cnt, off

    def infection_tables(self):
        fac = np.zeros(2, np.float32)
        shed = np.zeros(60, np.float32)
        self.L.or_get_infection_tables(self.h, fac.ctypes.data, shed.ctypes.data)
        return fac, shed.reshape(6, 10)

    def generation_offsets(self):
        md = self.params.max_days
        cp = np.zeros(md + 1, np.int32)
        cs = np.zeros(md + 1, np.int32)
        self.L.or_generation_offsets(self.h, cp.ctypes.data, cs.ctypes.data, md + 1)
        return cp, cs

    def set_offsets(self, workplace_off, errand_off):
        self.L.or_set_offsets(self.h, workplace_off, errand_off)

    def recalc_workplace(self, idx, age):
        return int(self.L.or_recalc_workplace(self.h, idx, age))

    def recalc_weekday(self, idx, age):
        d = (C.c_int32 * 2)()
        prof = self.L.or_recalc_weekday(self.h, idx, age, d)
        return int(d[0]), int(d[1]), int(prof)

    def recalc_weekend(self, idx):
        d = (C.c_int32 * 3)()
        self.L.or_recalc_weekend(self.h, idx, d)
        return int(d[0]), int(d[1]), int(d[2])

    def fish_workplace(self, r):
        return int(self.L.or_fish_workplace(self.h, r))

    def fish_weekday_dest(self, r, age):
        L = self.L
        L.or_fish_weekday_dest.restype = C.c_int32
        L.or_fish_weekday_dest.argtypes = [C.c_void_p, C.c_uint32, C.c_uint8]
        return int(L.or_fish_weekday_dest(self.h, r, age))

    def fish_weekend_dest(self, r):
        L = self.L
        L.or_fish_weekend_dest.restype = C.c_int32
        L.or_fish_weekend_dest.argtypes = [C.c_void_p, C.c_uint32]
        return int(L.or_fish_weekend_dest(self.h, r))

    def fish_school_and_age(self, r):
        age = C.c_uint8(0)
        school = C.c_int32(0)
        self.L.or_fish_school_and_age(self.h, r, C.byref(age), C.byref(school))
        return int(age.value), int(school.value)

    def enable_trace(self, on=True):
        self.L.or_enable_trace(self.h, 1 if on else 0)

    def trace(self):
        n = int(self.L.or_trace_size(self.h))
        inf = np.zeros(n, np.int32)
        vic = np.zeros(n, np.int32)
        typ = np.zeros(n, np.uint8)
        act = np.zeros(n, np.uint8)
        if n:
            self.L.or_get_trace(self.h, inf.ctypes.data, vic.ctypes.data, typ.ctypes.data, act.ctypes.data)
        return inf, vic, typ, act

    def set_state(self, day, rand_offset, arrays):
        """Continue from a day boundary: the six uint8 state arrays, the day
        about to run and the RNG offset reached by then."""
        a = [np.ascontiguousarray(x, np.uint8) for x in arrays]
        rc = self.L.or_set_state(self.h, day, rand_offset, *[x.ctypes.data for x in a])
        if rc != 0:
            raise RuntimeError("or_set_state before setup")

    def trace_draws(self):
        """Per slot of the last contact day: draw bits (2 pandemic, 1 seasonal
        succeeded) and profile draws (low nibble pandemic, high nibble seasonal)."""
        n = int(self.L.or_trace_size(self.h))
        draw = np.zeros(n, np.uint8)
        prof = np.zeros(n, np.uint8)
        if n:
            self.L.or_get_trace_draws(self.h, draw.ctypes.data, prof.ctypes.data)
        return draw, prof

    def infection_profile(self, r):
        return int(self.L.or_infection_profile(self.h, r))

    def infection_probability(self, profile, dd, strain, kval_sum):
        return float(self.L.or_infection_probability(self.h, profile, dd, strain, kval_sum))

    def write_outputs(self, directory, counts, infected):
        counts = np.ascontiguousarray(counts, np.int32)
        infected = np.ascontiguousarray(infected, np.int32)
        rc = self.L.or_write_outputs(self.h, directory.encode(), counts.ctypes.data, infected.ctypes.data,
                                    len(infected))
        assert rc == 0


def raw_table(name):
    L = lib()
    L.or_get_raw_table.argtypes = [C.c_char_p, C.c_void_p, C.c_int]
    out = np.zeros(128, np.float64)
    n = L.or_get_raw_table(name.encode(), out.ctypes.data, 128)
    if n < 0:
        raise KeyError(name)
    return out[:n].copy()


def fish_household_type(r):
    return int(lib().or_fish_household_type(C.c_uint32(r)))
