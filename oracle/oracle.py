"""ctypes wrapper of the oracle libraries (see oracle.c for what each follows).

Test infrastructure only: used as the checker by tests/, the smoke check and
the CPU-baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

from g4go_amd import abi

ORACLE_DIR = Path(__file__).resolve().parent
LIB32 = ORACLE_DIR / "_build" / "liboracle32.so"
LIB64 = ORACLE_DIR / "_build" / "liboracle64.so"
LIBREF = ORACLE_DIR / "_ref" / "libref_boundary.so"

_libs: dict = {}


def _load(precision: int) -> C.CDLL:
    if precision in _libs:
        return _libs[precision]
    path = LIB32 if precision == 32 else LIB64
    if not path.is_file():
        raise RuntimeError(f"{path} is missing: build it with `python -m oracle.build`")
    lib = C.CDLL(str(path))
    p = f"oracle{precision}_"
    real = C.c_float if precision == 32 else C.c_double
    getattr(lib, p + "philox").argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    getattr(lib, p + "uniform").argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]
    getattr(lib, p + "uniform").restype = C.c_float
    getattr(lib, p + "log").argtypes = [real]
    getattr(lib, p + "log").restype = real
    getattr(lib, p + "exp").argtypes = [real]
    getattr(lib, p + "exp").restype = real
    getattr(lib, p + "sincos_turn").argtypes = [real, C.POINTER(real), C.POINTER(real)]
    getattr(lib, p + "classify").argtypes = [real, real, C.c_int, C.c_int, real]
    getattr(lib, p + "absorption").argtypes = [real, C.c_int, real]
    getattr(lib, p + "fresnel").argtypes = [C.c_void_p, C.c_void_p]
    getattr(lib, p + "lambertian").argtypes = [C.c_void_p, real, real, C.c_void_p]
    getattr(lib, p + "find_surface").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_uint32]
    getattr(lib, p + "find_surface").restype = C.c_uint32
    getattr(lib, p + "run_synthetic").argtypes = [
        C.POINTER(abi.Scene), C.POINTER(abi.SyntheticSource), C.c_uint64, C.c_float, C.c_uint32, C.c_void_p,
        C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    getattr(lib, p + "run_synthetic").restype = C.c_int64
    getattr(lib, p + "run_emissions").argtypes = [
        C.POINTER(abi.Scene), C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint32, C.c_uint64, C.c_float, C.c_uint32,
        C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p]
    getattr(lib, p + "run_emissions").restype = C.c_int64
    getattr(lib, p + "rayleigh").argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32,
                                             C.c_void_p, C.c_void_p]
    getattr(lib, p + "box_exit").argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(real),
                                             C.POINTER(C.c_int)]
    getattr(lib, p + "adjoint_synthetic").argtypes = [
        C.POINTER(abi.Scene), C.POINTER(abi.SyntheticSource), C.c_uint64, C.c_float, C.c_uint32, C.c_void_p,
        C.c_uint64, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p]
    getattr(lib, p + "adjoint_emissions").argtypes = [
        C.POINTER(abi.Scene), C.c_void_p, C.c_uint64, C.c_uint64, C.c_float, C.c_uint32, C.c_void_p, C.c_uint64,
        C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p]
    getattr(lib, p + "mesh_trace").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_void_p, C.c_void_p,
                                               C.c_uint64, real, C.c_int, C.c_void_p, C.c_void_p]
    getattr(lib, p + "mesh_planes").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_void_p, C.c_uint32]
    getattr(lib, p + "mesh_planes").restype = C.c_uint32
    getattr(lib, p + "solid_kind").argtypes = [C.POINTER(abi.Scene), C.c_uint32]
    getattr(lib, p + "solid_contains").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_void_p, C.c_uint64, C.c_void_p]
    getattr(lib, p + "generate_emission").argtypes = [C.POINTER(abi.Scene), C.c_void_p, C.c_uint64, C.c_uint32,
                                                      C.c_uint32, C.c_void_p]
    getattr(lib, p + "generate_synthetic").argtypes = [C.POINTER(abi.Scene), C.POINTER(abi.SyntheticSource),
                                                       C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p]
    getattr(lib, p + "surface_reflect").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_void_p, C.c_void_p, real,
                                                    C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                    C.c_void_p, C.c_void_p, C.c_void_p]
    getattr(lib, p + "step_from").argtypes = [C.POINTER(abi.Scene), C.c_uint32, C.c_void_p, C.c_uint64, C.c_uint32,
                                              C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_void_p, C.c_void_p]
    for name in ("project_polarization", "reflect_direction"):
        getattr(lib, p + name).argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    getattr(lib, p + "normalize").argtypes = [C.c_void_p, C.c_void_p]
    getattr(lib, p + "reflect_polarization").argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    _libs[precision] = lib
    return lib


def _fn(precision: int, name: str):
    return getattr(_load(precision), f"oracle{precision}_{name}")


def _real(precision: int):
    return np.float32 if precision == 32 else np.float64


# ---- elementary pieces ----
def philox(ctr, key) -> np.ndarray:
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, np.uint32)
    _fn(32, "philox")(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    return out


def uniform(seed: int, photon_key: int, bounce: int, process: int, draw: int) -> float:
    return _fn(32, "uniform")(seed, photon_key, bounce, process, draw)


def log(x: float, precision: int = 32) -> float:
    return _fn(precision, "log")(x)


def exp(x: float, precision: int = 32) -> float:
    return _fn(precision, "exp")(x)


def sincos_turn(u: float, precision: int = 32) -> tuple[float, float]:
    real = C.c_float if precision == 32 else C.c_double
    s, c = real(), real()
    _fn(precision, "sincos_turn")(u, C.byref(s), C.byref(c))
    return s.value, c.value


def classify(r, t, has_r, has_t, sample, precision: int = 32) -> int:
    return _fn(precision, "classify")(r, t, int(has_r), int(has_t), sample)


def absorption(eff, sensor, sample, precision: int = 32) -> int:
    return _fn(precision, "absorption")(eff, int(sensor), sample)


def fresnel(cases: np.ndarray, precision: int = 32) -> np.ndarray:
    real = _real(precision)
    cases = np.ascontiguousarray(cases, dtype=real).reshape(-1, 11)
    out = np.zeros((cases.shape[0], 14), dtype=real)
    f = _fn(precision, "fresnel")
    for i in range(cases.shape[0]):
        f(cases[i].ctypes.data, out[i].ctypes.data)
    return out


def lambertian(normal, radial, azimuth, precision: int = 32) -> np.ndarray:
    real = _real(precision)
    n = np.ascontiguousarray(normal, dtype=real)
    out = np.zeros(3, real)
    _fn(precision, "lambertian")(n.ctypes.data, radial, azimuth, out.ctypes.data)
    return out


def vec_op(name: str, *vectors, precision: int = 32) -> np.ndarray:
    """normalize(a), project_polarization(pol, dir), reflect_direction(dir, n), reflect_polarization(pol, n, dir)."""
    real = _real(precision)
    args = [np.ascontiguousarray(v, dtype=real) for v in vectors]
    out = np.zeros(3, real)
    _fn(precision, name)(*[a.ctypes.data for a in args], out.ctypes.data)
    return out


def rayleigh(direction, polarization, seed, photon, event=0, step=0, precision: int = 64):
    real = _real(precision)
    d = np.ascontiguousarray(direction, dtype=real)
    p = np.ascontiguousarray(polarization, dtype=real)
    od, op = np.zeros(3, real), np.zeros(3, real)
    _fn(precision, "rayleigh")(d.ctypes.data, p.ctypes.data, seed, photon, event, step, od.ctypes.data,
                               op.ctypes.data)
    return od, op


def box_exit(lo, hi, p, d, precision: int = 32):
    real = _real(precision)
    lo = np.ascontiguousarray(lo, dtype=np.float32)
    hi = np.ascontiguousarray(hi, dtype=np.float32)
    pp = np.ascontiguousarray(p, dtype=real)
    dd = np.ascontiguousarray(d, dtype=real)
    dist = (C.c_float if precision == 32 else C.c_double)()
    axis = C.c_int()
    _fn(precision, "box_exit")(lo.ctypes.data, hi.ctypes.data, pp.ctypes.data, dd.ctypes.data, C.byref(dist),
                               C.byref(axis))
    return dist.value, axis.value


def mesh_trace(scene, volume_id: int, p, d, tmin: float = 0.0, entry: bool = False, precision: int = 32):
    """Exit (entry=False) or entry distances of rays into a mesh volume, and the face crossed."""
    real = _real(precision)
    p = np.ascontiguousarray(p, dtype=real).reshape(-1, 3)
    d = np.ascontiguousarray(d, dtype=real).reshape(-1, 3)
    dist = np.zeros(len(p), real)
    face = np.zeros(len(p), np.uint32)
    desc, keep = scene.to_abi()
    _fn(precision, "mesh_trace")(C.byref(desc), volume_id, p.ctypes.data, d.ctypes.data, len(p), tmin, int(entry),
                                 dist.ctypes.data, face.ctypes.data)
    return dist, face


def mesh_planes(scene, volume_id: int, precision: int = 32) -> np.ndarray:
    """Face planes of a mesh volume as rows (n[3], h, s[3])."""
    real = _real(precision)
    out = np.zeros((1024, 7), real)
    desc, keep = scene.to_abi()
    n = _fn(precision, "mesh_planes")(C.byref(desc), volume_id, out.ctypes.data, 1024)
    return out[:n].copy()


def solid_kind(scene, volume_id: int) -> int:
    """0 box, 1 convex mesh (face planes), 2 triangle solid, 3 open triangle solid (no inside)."""
    desc, keep = scene.to_abi()
    return _fn(32, "solid_kind")(C.byref(desc), volume_id)


def solid_contains(scene, volume_id: int, q, precision: int = 32) -> np.ndarray:
    """Whether each point lies inside the volume's solid."""
    real = _real(precision)
    q = np.ascontiguousarray(q, dtype=real).reshape(-1, 3)
    out = np.zeros(len(q), np.int32)
    desc, keep = scene.to_abi()
    _fn(precision, "solid_contains")(C.byref(desc), volume_id, q.ctypes.data, len(q), out.ctypes.data)
    return out.astype(bool)


def find_surface(scene, from_id: int, to_id: int) -> int:
    desc, keep = scene.to_abi()
    r = _fn(32, "find_surface")(C.byref(desc), from_id, to_id)
    del keep
    return r


# ---- generation, surfaces and single steps (distribution tests) ----
STATE_FIELDS = ("x", "y", "z", "dx", "dy", "dz", "px", "py", "pz", "energy", "time", "valid")


def generate_emission(scene, record: np.ndarray, seed: int, n: int, first: int = 0, precision: int = 32) -> np.ndarray:
    """Photons first .. first + n - 1 of one emission record, as (n, 12) rows of STATE_FIELDS."""
    desc, keep = scene.to_abi()
    rec = np.ascontiguousarray(np.asarray(record, dtype=abi.EMISSION_DTYPE).reshape(1))
    out = np.zeros((n, 12), np.float64)
    _fn(precision, "generate_emission")(C.byref(desc), rec.ctypes.data, seed, first, n, out.ctypes.data)
    return out


def generate_synthetic(scene, source: abi.SyntheticSource, seed: int, event: int, n: int, first: int = 0,
                       precision: int = 32) -> np.ndarray:
    """Photons first .. first + n - 1 of one synthetic event, as (n, 12) rows of STATE_FIELDS."""
    desc, keep = scene.to_abi()
    out = np.zeros((n, 12), np.float64)
    _fn(precision, "generate_synthetic")(C.byref(desc), C.byref(source), seed, event, first, n, out.ctypes.data)
    return out


def surface_reflect(scene, surface: int, direction, normal, energy_ev: float, seed: int, n: int, event: int = 0,
                    step: int = 0, first: int = 0, precision: int = 32):
    """Reflections off surface index `surface` of n photons arriving along `direction` at a face with
    `normal` (facing the photon): (directions (n, 3), facet normals (n, 3), unified branch (n,))."""
    real = _real(precision)
    desc, keep = scene.to_abi()
    d = np.ascontiguousarray(direction, dtype=real)
    nn = np.ascontiguousarray(normal, dtype=real)
    out_d = np.zeros((n, 3), real)
    out_f = np.zeros((n, 3), real)
    branch = np.zeros(n, np.int32)
    _fn(precision, "surface_reflect")(C.byref(desc), surface, d.ctypes.data, nn.ctypes.data, energy_ev, seed, event,
                                      step, first, n, out_d.ctypes.data, out_f.ctypes.data, branch.ctypes.data)
    return out_d, out_f, branch


def step_from(scene, volume_id: int, state, seed: int, n: int, event: int = 0, step: int = 0, first: int = 0,
              eps: float = 1e-4, precision: int = 32):
    """One transport step of photons first .. + n - 1 from one state (pos, dir, pol, energy, time): the
    states after the step ((n, 12); column 11 = volume index) and the step statuses."""
    desc, keep = scene.to_abi()
    st = np.zeros(12, np.float64)
    st[:11] = np.asarray(state, dtype=np.float64)[:11]
    out = np.zeros((n, 12), np.float64)
    status = np.zeros(n, np.int32)
    _fn(precision, "step_from")(C.byref(desc), volume_id, st.ctypes.data, seed, event, step, first, n, eps,
                                out.ctypes.data, status.ctypes.data)
    return out, status


# ---- whole-path runs ----
class OracleResult:
    def __init__(self, hits: np.ndarray, stats: np.ndarray, hist: np.ndarray) -> None:
        self.hits = hits        # abi.HIT_DTYPE, in launch order (the order of the batch's photons)
        self.stats = stats      # (n_events, 9): detected absorbed escaped truncated max_bounce invalid zero steps generated
        self.hist = hist        # 258 slots

    STAT_NAMES = ("detected", "absorbed", "escaped", "truncated", "max_bounce", "invalid", "zero_step", "steps",
                  "generated")


def run_synthetic(scene, source: abi.SyntheticSource, seed: int, event_ids, photon_counts, precision: int = 32,
                  eps: float = 1e-4, max_steps: int = 4096, photon_begin: int = 0, photon_end: int = 0xFFFFFFFF,
                  threads: int = 1) -> OracleResult:
    """Transports the synthetic batch on the CPU; `threads` splits the photon range of each event."""
    desc, keep = scene.to_abi()
    ev = np.ascontiguousarray(event_ids, dtype=np.uint32)
    counts = np.ascontiguousarray(photon_counts, dtype=np.uint32)
    fn = _fn(precision, "run_synthetic")
    total = int(np.minimum(counts, photon_end).astype(np.int64).sum())

    def work(lo: int, hi: int):
        cap = max(1, total)
        hits = np.zeros(cap, dtype=abi.HIT_DTYPE)
        stats = np.zeros((ev.size, 9), np.uint64)
        hist = np.zeros(abi.HIST_SLOTS, np.uint64)
        n = fn(C.byref(desc), C.byref(source), seed, eps, max_steps, ev.ctypes.data, counts.ctypes.data, ev.size, lo,
               hi, hits.ctypes.data, cap, stats.ctypes.data, hist.ctypes.data)
        if n < 0:
            raise RuntimeError("oracle hit buffer overflow")
        return hits[:n], stats, hist

    end = min(int(counts.max()) if counts.size else 0, photon_end)
    if threads <= 1 or end - photon_begin < 2 * threads:
        parts = [work(photon_begin, end)]
    else:
        edges = np.linspace(photon_begin, end, threads + 1).astype(np.int64)
        with ThreadPoolExecutor(threads) as pool:
            parts = list(pool.map(lambda i: work(int(edges[i]), int(edges[i + 1])), range(threads)))
    hits = np.concatenate([p[0] for p in parts])
    # launch order: events in batch order, photons ascending
    position = {int(e): i for i, e in enumerate(ev)}
    pos = np.array([position[int(e)] for e in hits["event_id"]], np.int64) if hits.size else np.zeros(0, np.int64)
    hits = hits[np.lexsort((hits["photon_id"], pos))]
    stats = np.zeros((ev.size, 9), np.uint64)
    for p in parts:
        stats[:, [0, 1, 2, 3, 5, 6, 7, 8]] += p[1][:, [0, 1, 2, 3, 5, 6, 7, 8]]
        stats[:, 4] = np.maximum(stats[:, 4], p[1][:, 4])
    hist = sum(p[2] for p in parts)
    del keep
    return OracleResult(hits, stats, hist)


def run_emissions(scene, emissions: np.ndarray, event_ids, emission_event_indices, seed: int,
                  precision: int = 32, eps: float = 1e-4, max_steps: int = 4096, threads: int = 1) -> OracleResult:
    """Transports the emission batch on the CPU; `threads` splits the records (by photon count)."""
    desc, keep = scene.to_abi()
    em = np.ascontiguousarray(emissions, dtype=abi.EMISSION_DTYPE)
    idx = np.ascontiguousarray(emission_event_indices, dtype=np.uint32)
    n_events = len(event_ids)
    fn = _fn(precision, "run_emissions")

    def work(lo: int, hi: int):
        part, part_idx = em[lo:hi], idx[lo:hi]
        cap = max(1, int(part["photon_count"].astype(np.int64).sum()))
        hits = np.zeros(cap, dtype=abi.HIT_DTYPE)
        stats = np.zeros((n_events, 9), np.uint64)
        hist = np.zeros(abi.HIST_SLOTS, np.uint64)
        n = fn(C.byref(desc), part.ctypes.data, part.size, part_idx.ctypes.data, n_events, seed, eps, max_steps,
               hits.ctypes.data, cap, stats.ctypes.data, hist.ctypes.data)
        if n < 0:
            raise RuntimeError(f"oracle run failed ({n})")
        return hits[:n], stats, hist

    if threads <= 1 or em.size < 2 * threads:
        parts = [work(0, em.size)]
    else:
        cum = np.cumsum(em["photon_count"].astype(np.int64))
        cuts = np.searchsorted(cum, np.linspace(0, cum[-1], threads + 1)[1:-1])
        edges = [0, *[int(c) for c in cuts], em.size]
        with ThreadPoolExecutor(threads) as pool:
            parts = list(pool.map(lambda i: work(edges[i], edges[i + 1]), range(threads)))
    hits = np.concatenate([p[0] for p in parts])  # launch order: records in order, photons ascending
    stats = np.zeros((n_events, 9), np.uint64)
    for p in parts:
        stats[:, [0, 1, 2, 3, 5, 6, 7, 8]] += p[1][:, [0, 1, 2, 3, 5, 6, 7, 8]]
        stats[:, 4] = np.maximum(stats[:, 4], p[1][:, 4])
    hist = sum(p[2] for p in parts)
    del keep
    return OracleResult(hits, stats, hist)


def _adjoint(call, hits, slot_weight, tables, n_params: int, threads: int):
    h = np.ascontiguousarray(hits, dtype=abi.HIT_DTYPE)
    w = np.ascontiguousarray(slot_weight, dtype=np.float64)
    t = np.ascontiguousarray(np.asarray(tables, dtype=np.uint32).reshape(-1, 2))

    def work(lo: int, hi: int):
        part = h[lo:hi]
        grad = np.zeros(n_params, np.float64)
        counters = np.zeros(4, np.uint64)
        rc = call(part.ctypes.data if part.size else None, part.size, w.ctypes.data, t.ctypes.data, t.shape[0],
                  grad.ctypes.data, n_params, counters.ctypes.data)
        if rc != 0:
            raise ValueError("parameter layout does not match n_params")
        return grad, counters

    if threads <= 1 or h.size < 2 * threads:
        parts = [work(0, h.size)]
    else:
        edges = np.linspace(0, h.size, threads + 1).astype(np.int64)
        with ThreadPoolExecutor(threads) as pool:
            parts = list(pool.map(lambda i: work(int(edges[i]), int(edges[i + 1])), range(threads)))
    return sum(p[0] for p in parts), sum(p[1] for p in parts)


def adjoint_synthetic(scene, source, seed: int, hits: np.ndarray, slot_weight, tables, n_params: int,
                      precision: int = 32, eps: float = 1e-4, max_steps: int = 4096, threads: int = 1):
    """Replay gradient over `hits` of a synthetic batch (FP64 accumulation); returns (grad, counters)."""
    desc, keep = scene.to_abi()
    fn = _fn(precision, "adjoint_synthetic")
    out = _adjoint(lambda *a: fn(C.byref(desc), C.byref(source), seed, eps, max_steps, *a), hits, slot_weight,
                   tables, n_params, threads)
    del keep
    return out


def adjoint_emissions(scene, emissions: np.ndarray, seed: int, hits: np.ndarray, slot_weight, tables, n_params: int,
                      precision: int = 32, eps: float = 1e-4, max_steps: int = 4096, threads: int = 1):
    """Replay gradient over `hits` of an emission-record batch; returns (grad, counters)."""
    desc, keep = scene.to_abi()
    em = np.ascontiguousarray(emissions, dtype=abi.EMISSION_DTYPE)
    fn = _fn(precision, "adjoint_emissions")
    out = _adjoint(lambda *a: fn(C.byref(desc), em.ctypes.data, em.size, seed, eps, max_steps, *a), hits,
                   slot_weight, tables, n_params, threads)
    del keep
    return out


# ---- the reference's own boundary physics, compiled from its header ----
def reference_library():
    """libref_boundary.so or None when it has not been built."""
    if "ref" in _libs:
        return _libs["ref"]
    if not LIBREF.is_file():
        return None
    lib = C.CDLL(str(LIBREF))
    lib.ref_compute_fresnel.argtypes = [C.c_void_p, C.c_void_p]
    lib.ref_project_polarization.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ref_reflect_direction.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ref_reflect_polarization.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lib.ref_classify_surface.argtypes = [C.c_float, C.c_float, C.c_int, C.c_int, C.c_float]
    lib.ref_evaluate_absorption.argtypes = [C.c_float, C.c_int, C.c_float]
    lib.ref_sample_lambertian.argtypes = [C.c_void_p, C.c_float, C.c_float, C.c_void_p]
    _libs["ref"] = lib
    return lib
