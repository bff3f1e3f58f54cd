"""Golden cases written by tests/golden/make_golden.py from the reference.

Each case rebuilds the circuit with pyjjasim_amd and returns the keyword
arguments of TimeEvolutionProblem (the same inputs the reference was run with)
and the reference outputs.
"""
import os

import numpy as np
import scipy.sparse

import pyjjasim_amd as pa

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
_traj = None
_graphs = None


def traj():
    global _traj
    if _traj is None:
        _traj = np.load(os.path.join(GOLDEN, "trajectories.npz"), allow_pickle=False)
    return _traj


def graphs():
    global _graphs
    if _graphs is None:
        _graphs = np.load(os.path.join(GOLDEN, "graphs.npz"), allow_pickle=False)
    return _graphs


def g(case, key):
    return traj()[f"{case}__{key}"]


def build(case):
    """-> (circuit, kwargs for TimeEvolutionProblem, seed or None)"""
    seed = None
    if case == "c1":
        c = pa.SquareArray(10, 10)
        kw = dict(time_step=0.05, time_step_count=10000, external_flux=0.1,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "sq6_w4_c1":
        c = pa.SquareArray(6, 6)
        c.set_capacitance(1.0)
        kw = dict(time_step=0.05, time_step_count=800, external_flux=0.25,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "hc6":
        c = pa.HoneycombArray(6, 6)
        kw = dict(time_step=0.05, time_step_count=1000, external_flux=0.05,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "tri6":
        c = pa.TriangularArray(6, 6)
        kw = dict(time_step=0.05, time_step_count=600, external_flux=0.15,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "sq8_L":
        c = pa.SquareArray(8, 8)
        c.set_inductance(1.0)
        c.set_capacitance(1.0)
        kw = dict(time_step=0.05, time_step_count=1000, external_flux=0.1,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "sq5_mutual":
        c = pa.SquareArray(5, 5)
        Nj = c.junction_count()
        L = scipy.sparse.csc_matrix((g(case, "in_L_data"), g(case, "in_L_indices"),
                                     g(case, "in_L_indptr")), shape=(Nj, Nj))
        c.set_inductance(L)
        kw = dict(time_step=0.05, time_step_count=500, external_flux=0.2,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "sq5_vs":
        c = pa.SquareArray(5, 5)
        kw = dict(time_step=0.05, time_step_count=400, voltage_sources=g(case, "in_Vs"),
                  store_time_steps=g(case, "in_ts"))
    elif case in ("sq5_s4", "sq5_s5"):
        c = pa.SquareArray(5, 5)
        c.set_capacitance(0.5)
        kw = dict(time_step=0.05, time_step_count=500, external_flux=0.1,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"),
                  stencil_width=int(case[-1]))
    elif case == "sq6_tdep":
        c = pa.SquareArray(6, 6)
        kw = dict(time_step=0.05, time_step_count=400, external_flux=g(case, "in_f"),
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"))
    elif case == "sq76_hetero":
        c = pa.SquareArray(7, 6)
        c.set_critical_current(g(case, "in_Ic"))
        c.set_resistance(g(case, "in_R"))
        c.set_capacitance(g(case, "in_C"))
        kw = dict(time_step=0.05, time_step_count=500, external_flux=g(case, "in_f"),
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"),
                  config_at_minus_1=g(case, "in_th1"), config_at_minus_2=g(case, "in_th2"))
    elif case == "sq6_T":
        c = pa.SquareArray(6, 6)
        kw = dict(time_step=0.05, time_step_count=200, external_flux=0.1,
                  current_sources=g(case, "in_Is"), temperature=0.05,
                  store_time_steps=g(case, "in_ts"))
        seed = int(g(case, "in_seed"))
    elif case == "sq17_T":
        c = pa.SquareArray(17, 17)
        kw = dict(time_step=0.05, time_step_count=60, external_flux=0.2,
                  current_sources=g(case, "in_Is"), temperature=0.02,
                  store_time_steps=g(case, "in_ts"))
        seed = int(g(case, "in_seed"))
    elif case == "sq9_iv":
        c = pa.SquareArray(9, 9)
        c.set_capacitance(1.0)
        kw = dict(time_step=0.05, time_step_count=600, external_flux=0.25,
                  current_sources=g(case, "in_Is"), store_time_steps=g(case, "in_ts"),
                  store_current=False, store_voltage=False)
    else:
        raise KeyError(case)
    return c, kw, seed


CASES = ["c1", "sq6_w4_c1", "hc6", "tri6", "sq8_L", "sq5_mutual", "sq5_vs", "sq5_s4", "sq5_s5",
         "sq6_tdep", "sq76_hetero", "sq6_T", "sq17_T", "sq9_iv"]
T_CASES = ["sq6_T", "sq17_T"]


def oracle_kwargs(kw):
    """Translate TimeEvolutionProblem kwargs into oracle.te_oracle.evolve kwargs."""
    m = dict(dt=kw["time_step"], Nt=kw["time_step_count"], f=kw.get("external_flux", 0.0),
             Is=kw.get("current_sources", 0.0), Vs=kw.get("voltage_sources", 0.0),
             T=kw.get("temperature", 0.0), store_time_steps=kw.get("store_time_steps"),
             store_theta=kw.get("store_theta", True), store_voltage=kw.get("store_voltage", True),
             store_current=kw.get("store_current", True), stencil_width=kw.get("stencil_width", 3))
    cfg = [kw[k] for k in ("config_at_minus_1", "config_at_minus_2") if k in kw]
    if cfg:
        m["configs"] = cfg
    return m
