"""Output path after the transport: the per-event split of a batch with the
scheduler's invariants, the two ntuples G4GO writes per event, and the
CPU/GPU regression gates -- all without Geant4 or ROOT.

* split_by_event: OpticalBatchScheduler::CompleteBatch
  (src/optical/geant4/OpticalBatchScheduler.cpp:243-446): every event of the
  batch gets its statistics and detections; the output must hold one
  statistics record per requested event, no unknown or duplicate event, event
  counters that add up to the batch counters (maximum-bounce: the maximum), and
  per-event detected counts equal to the event's detections.
* CrystalHit / SensorHit rows: the ntuples of RunAction.cpp:29-43, filled as
  Analysis::WriteEvent does (Analysis.cpp:46-83): one CrystalHit row per
  crystal deposit (eventID, moduleID, Edep [MeV], nOptPho = the event's
  detected count, nGenOptPho) and one SensorHit row per event with detections
  (eventID, sensorID[], timeOfFlight[] as float).  write_ntuples stores them as
  Parquet (list columns for SensorHit).
* regression_gates: the gates of test/TestOpticalTransport.cxx:154-352 --
  histograms nOptPho (100 bins over [0, 2000)), Edep (50 over [0, 6) MeV), time
  of flight (100 over [0, 1000) ns), sensor occupancy (64 over [0, 64)); the
  unweighted two-histogram chi-square test and the binned Kolmogorov test of
  ROOT's TH1 (Chi2Test "UU", KolmogorovTest); pass when the total detected
  photons and the sensor counts agree within 10%, nOptPho mean within 5% and RMS
  within 10%, Edep chi-square z <= 5 with entries within 1%, mean 2%, RMS 5%,
  time of flight Kolmogorov p >= 0.01 with mean within 1% and RMS 5%, occupancy
  chi-square z <= 5, and no value outside its histogram.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import special, stats

from . import abi

COUNTERS = ("detected", "absorbed", "escaped", "truncated", "invalid", "zero_step")


def split_by_event(output, event_ids) -> dict:
    """{event_id: (event_statistics record, detections)} of a completed batch, checked
    as OpticalBatchScheduler::CompleteBatch checks it (raises RuntimeError)."""
    event_ids = [int(e) for e in event_ids]
    index = {e: i for i, e in enumerate(event_ids)}
    ev = output.event_statistics
    if len(ev) != len(event_ids):
        raise RuntimeError("output event statistics count does not match the batch")
    seen = [False] * len(event_ids)
    for rec in ev:
        i = index.get(int(rec["event_id"]))
        if i is None:
            raise RuntimeError("output contains an unknown event ID")
        if seen[i]:
            raise RuntimeError("output contains duplicate event statistics")
        seen[i] = True
    st = output.statistics
    for name in COUNTERS:
        if int(ev[name].astype(np.uint64).sum()) != int(st[name]):
            raise RuntimeError(f"batch/event statistics mismatch in {name} count")
    if len(ev) and int(ev["max_bounce"].max()) != int(st["max_bounce"]):
        raise RuntimeError("batch/event statistics mismatch in maximum-bounce count")
    det = output.detections
    per_event = {e: [] for e in event_ids}
    if len(det):
        order = np.argsort(det["event_id"], kind="stable")
        ids, starts = np.unique(det["event_id"][order], return_index=True)
        bounds = list(starts) + [len(order)]
        for k, e in enumerate(ids):
            if int(e) not in index:
                raise RuntimeError("output contains a detection for an unknown event ID")
            per_event[int(e)] = det[order[bounds[k]:bounds[k + 1]]]
    result = {}
    for rec in ev:
        e = int(rec["event_id"])
        hits = per_event[e] if len(per_event[e]) else np.zeros(0, abi.HIT_DTYPE)
        if int(rec["detected"]) != len(hits):
            raise RuntimeError("event detected count does not match compacted hits")
        result[e] = (rec, hits)
    return result


@dataclass
class Ntuples:
    """Columns of the CrystalHit and SensorHit ntuples."""
    crystal: dict = field(default_factory=dict)   # eventID, moduleID, Edep, nOptPho, nGenOptPho (arrays)
    sensor_event: np.ndarray = None                # eventID per SensorHit row
    sensor_id: list = None                         # per row: int32 array
    time_of_flight: list = None                    # per row: float32 array


def ntuples(events: dict, module_of_event: dict, edep_mev_of_event: dict) -> Ntuples:
    """Ntuple rows from split_by_event's result; one crystal deposit per event
    (module and energy given per event)."""
    ids = sorted(events)
    rows = {"eventID": [], "moduleID": [], "Edep": [], "nOptPho": [], "nGenOptPho": []}
    s_event, s_id, s_tof = [], [], []
    for e in ids:
        rec, hits = events[e]
        rows["eventID"].append(e)
        rows["moduleID"].append(int(module_of_event[e]))
        rows["Edep"].append(float(edep_mev_of_event[e]))
        rows["nOptPho"].append(int(rec["detected"]))
        rows["nGenOptPho"].append(int(rec["captured"]))
        if len(hits):
            s_event.append(e)
            s_id.append(hits["sensor_id"].astype(np.int32))
            s_tof.append(hits["time_ns"].astype(np.float32))
    crystal = {"eventID": np.array(rows["eventID"], np.int32), "moduleID": np.array(rows["moduleID"], np.int32),
               "Edep": np.array(rows["Edep"], np.float64), "nOptPho": np.array(rows["nOptPho"], np.int32),
               "nGenOptPho": np.array(rows["nGenOptPho"], np.int32)}
    return Ntuples(crystal, np.array(s_event, np.int32), s_id, s_tof)


def write_ntuples(nt: Ntuples, path_prefix: str) -> tuple[str, str]:
    """<prefix>_CrystalHit.parquet and <prefix>_SensorHit.parquet."""
    import pyarrow as pa
    import pyarrow.parquet as pq

    a = f"{path_prefix}_CrystalHit.parquet"
    b = f"{path_prefix}_SensorHit.parquet"
    pq.write_table(pa.table(nt.crystal), a)
    pq.write_table(pa.table({"eventID": nt.sensor_event,
                             "sensorID": pa.array(nt.sensor_id, type=pa.list_(pa.int32())),
                             "timeOfFlight": pa.array(nt.time_of_flight, type=pa.list_(pa.float32()))}), b)
    return a, b


# ---- ROOT TH1 statistics, restated ----
def histogram(values, bins: int, lo: float, hi: float) -> np.ndarray:
    """Counts with ROOT's layout: [underflow, bins..., overflow]."""
    v = np.asarray(values, np.float64)
    h = np.zeros(bins + 2, np.float64)
    h[0] = np.count_nonzero(v < lo)
    h[-1] = np.count_nonzero(v >= hi)
    inside = v[(v >= lo) & (v < hi)]
    idx = np.minimum(((inside - lo) * (bins / (hi - lo))).astype(np.int64), bins - 1)
    h[1:-1] = np.bincount(idx, minlength=bins)
    return h


def chi2_test_uu(h1: np.ndarray, h2: np.ndarray, overflow: bool = True) -> float:
    """p-value of ROOT's TH1::Chi2Test for two unweighted histograms
    (Gagunashvili): chi2 = sum (N2 n1 - N1 n2)^2 / (N1 N2 (n1 + n2)) over bins
    with entries, ndf = bins used - 1."""
    sl = slice(None) if overflow else slice(1, -1)
    a, b = h1[sl].astype(np.float64), h2[sl].astype(np.float64)
    n1, n2 = a.sum(), b.sum()
    used = (a + b) > 0
    if n1 == 0 or n2 == 0 or used.sum() < 2:
        return float("nan")
    chi2 = float(np.sum((n2 * a[used] - n1 * b[used]) ** 2 / (a[used] + b[used])) / (n1 * n2))
    ndf = int(used.sum()) - 1
    return float(stats.chi2.sf(chi2, ndf))


def kolmogorov_test(h1: np.ndarray, h2: np.ndarray) -> float:
    """ROOT's binned TH1::KolmogorovTest probability (in-range bins): the largest
    difference of the normalised cumulative contents, z = d sqrt(n1 n2 / (n1 + n2)),
    p = KolmogorovProb(z) = 2 sum_k (-1)^(k-1) exp(-2 k^2 z^2)."""
    a, b = h1[1:-1].astype(np.float64), h2[1:-1].astype(np.float64)
    n1, n2 = a.sum(), b.sum()
    if n1 == 0 or n2 == 0:
        return float("nan")
    d = float(np.max(np.abs(np.cumsum(a) / n1 - np.cumsum(b) / n2)))
    z = d * np.sqrt(n1 * n2 / (n1 + n2))
    return float(special.kolmogorov(z))


def z_score(p: float) -> float:
    return float("inf") if p == 0.0 else float(-special.ndtri(p / 2.0))


def _rel(a: float, b: float) -> float:
    return abs(a - b) / max(abs(b), np.finfo(np.float64).eps)


def regression_gates(cpu: Ntuples, gpu: Ntuples, n_opt_range=(100, 0.0, 2000.0), edep_range=(50, 0.0, 6.0),
                     tof_range=(100, 0.0, 1000.0), sensors: int = 64) -> dict:
    """The CPU/GPU regression of test/TestOpticalTransport.cxx on two sets of ntuples."""
    r = {}
    for tag, nt in (("cpu", cpu), ("gpu", gpu)):
        n = nt.crystal["nOptPho"].astype(np.float64)
        e = nt.crystal["Edep"].astype(np.float64)
        ids = np.concatenate(nt.sensor_id) if nt.sensor_id else np.zeros(0, np.int32)
        tof = np.concatenate(nt.time_of_flight).astype(np.float64) if nt.time_of_flight else np.zeros(0)
        if len(n) < 500:
            raise RuntimeError("CrystalHit contains too few entries")
        if len(ids) == 0:
            raise RuntimeError("SensorHit contains no detections")
        r[tag] = {
            "entries": len(n), "total": int(n.sum()), "mean": float(n.mean()), "rms": float(n.std()),
            "edep_entries": len(e), "edep_mean": float(e.mean()), "edep_rms": float(e.std()),
            "sensor_entries": len(ids), "tof_mean": float(tof.mean()), "tof_rms": float(tof.std()),
            "h_nopt": histogram(n, *n_opt_range), "h_edep": histogram(e, *edep_range),
            "h_tof": histogram(tof, *tof_range), "h_occ": histogram(ids, sensors, 0.0, float(sensors)),
            "invalid_nopt": int(np.count_nonzero(n < 0)), "invalid_edep": int(np.count_nonzero(~np.isfinite(e) | (e < 0))),
            "invalid_sensor_rows": sum(1 for s, t in zip(nt.sensor_id, nt.time_of_flight)
                                       if len(s) == 0 or len(s) != len(t) or np.any(s < 0) or np.any(s >= sensors)
                                       or not np.all(np.isfinite(t)) or np.any(t < 0)),
        }
    c, g = r["cpu"], r["gpu"]
    if c["total"] == 0 or g["total"] == 0:
        raise RuntimeError("nOptPho total is zero")
    oor = lambda h: h[0] != 0 or h[-1] != 0  # noqa: E731
    nopt_valid = c["invalid_nopt"] == 0 and g["invalid_nopt"] == 0 and not oor(c["h_nopt"]) and not oor(g["h_nopt"])
    edep_valid = c["invalid_edep"] == 0 and g["invalid_edep"] == 0 and not oor(c["h_edep"]) and not oor(g["h_edep"])
    sens_ok = c["invalid_sensor_rows"] == 0 and g["invalid_sensor_rows"] == 0
    tof_valid = sens_ok and not oor(c["h_tof"]) and not oor(g["h_tof"])
    occ_valid = sens_ok and not oor(c["h_occ"]) and not oor(g["h_occ"])
    p = {"nopt": chi2_test_uu(c["h_nopt"], g["h_nopt"]), "tof": chi2_test_uu(c["h_tof"], g["h_tof"]),
         "tof_shape": kolmogorov_test(c["h_tof"], g["h_tof"]), "occ": chi2_test_uu(c["h_occ"], g["h_occ"]),
         "edep": chi2_test_uu(c["h_edep"], g["h_edep"])}

    def passed(pv, valid):
        return bool(valid and np.isfinite(pv) and 0.0 <= pv <= 1.0 and z_score(pv) <= 5.0)

    out = {
        "total_rel_diff": _rel(g["total"], c["total"]), "sensor_rel_diff": _rel(g["sensor_entries"], c["sensor_entries"]),
        "nopt_mean_rel_diff": _rel(g["mean"], c["mean"]), "nopt_rms_rel_diff": _rel(g["rms"], c["rms"]),
        "edep_entry_rel_diff": _rel(g["edep_entries"], c["edep_entries"]),
        "edep_mean_rel_diff": _rel(g["edep_mean"], c["edep_mean"]), "edep_rms_rel_diff": _rel(g["edep_rms"], c["edep_rms"]),
        "tof_mean_rel_diff": _rel(g["tof_mean"], c["tof_mean"]), "tof_rms_rel_diff": _rel(g["tof_rms"], c["tof_rms"]),
        "p_values": p, "z_scores": {k: z_score(v) if np.isfinite(v) else float("nan") for k, v in p.items()},
        "cpu": {k: v for k, v in c.items() if not k.startswith("h_")},
        "gpu": {k: v for k, v in g.items() if not k.startswith("h_")},
    }
    out["total_passed"] = out["total_rel_diff"] <= 0.10
    out["sensor_total_passed"] = out["sensor_rel_diff"] <= 0.10
    out["nopt_passed"] = bool(nopt_valid and out["nopt_mean_rel_diff"] <= 0.05 and out["nopt_rms_rel_diff"] <= 0.10)
    out["edep_passed"] = bool(passed(p["edep"], edep_valid) and out["edep_entry_rel_diff"] <= 0.01
                              and out["edep_mean_rel_diff"] <= 0.02 and out["edep_rms_rel_diff"] <= 0.05)
    out["tof_passed"] = bool(tof_valid and p["tof_shape"] >= 0.01 and out["tof_mean_rel_diff"] <= 0.01
                             and out["tof_rms_rel_diff"] <= 0.05)
    out["occupancy_passed"] = passed(p["occ"], occ_valid)
    out["passed"] = bool(out["total_passed"] and out["sensor_total_passed"] and out["nopt_passed"] and out["edep_passed"]
                         and out["tof_passed"] and out["occupancy_passed"])
    return out
