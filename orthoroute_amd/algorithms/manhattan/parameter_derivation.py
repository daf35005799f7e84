"""Negotiation schedule derived from board characteristics.

Reference: orthoroute/algorithms/manhattan/parameter_derivation.py:51-267
(derive_routing_parameters, apply_derived_parameters).  apply_derived_parameters
raises max_iterations to at least the derived 250 like the reference; the
router's `max_iterations_cap` attribute, when set, caps it afterwards (the
benchmark configurations quote iteration caps such as 30).
"""
from dataclasses import dataclass


@dataclass
class DerivedRoutingParameters:
    pres_fac_init: float
    pres_fac_mult: float
    pres_fac_max: float
    hist_cost_weight: float
    hist_gain: float
    history_decay: float
    hotset_cap: int
    hotset_percentage: float
    layer_bias_alpha: float
    layer_bias_min: float
    layer_bias_max: float
    via_cost_base: float
    stagnation_patience: int
    max_iterations: int
    strategy: str


def derive_routing_parameters(board) -> DerivedRoutingParameters:
    rho, L, N = board.congestion_ratio, board.layer_count, board.net_count
    if rho < 0.6:
        mult, pmax, strat = 1.15, 6.0, "SPARSE (fast convergence)"
    elif rho < 0.9:
        mult, pmax, strat = 1.12, 7.0, "NORMAL (balanced)"
    elif rho < 1.2:
        mult, pmax, strat = 1.10, 8.0, "TIGHT (gentle escalation)"
    else:
        mult, pmax, strat = 1.08, 10.0, "DENSE (conservative)"
    if L <= 12:
        pmax *= 0.75
    elif L >= 25:
        pmax *= 1.25
    if L <= 10:
        layer_bonus = 6.0
    elif L <= 15:
        layer_bonus = 4.0
    elif L <= 25:
        layer_bonus = 2.0
    else:
        layer_bonus = 0.0
    hist_w = 8.0 + layer_bonus + 6.0 * max(0.0, min(1.0, rho - 0.7))
    hist_gain = 0.20 + 0.15 * max(0.0, min(1.0, rho - 0.8))
    if rho < 0.7:
        hp = 0.25
    elif rho < 1.0:
        hp = 0.20
    elif rho < 1.3:
        hp = 0.15
    else:
        hp = 0.12
    hotset_cap = max(32, min(int(N * hp), int(N * 0.30)))
    if L <= 12:
        lb = (0.20, 0.60, 1.80)
    elif L <= 20:
        lb = (0.12, 0.75, 1.50)
    else:
        lb = (0.08, 0.85, 1.20)
    vf = board.via_flexibility
    via_cost = 0.5 if vf > 0.9 else (0.7 if vf > 0.7 else 1.0)
    if rho < 0.8:
        patience = 5
    elif rho < 1.1:
        patience = 6
    else:
        patience = 8
    max_it = 250
    if L <= 12:
        max_it = max(max_it, 250)
        patience = max(patience, 7)
    return DerivedRoutingParameters(
        pres_fac_init=1.0, pres_fac_mult=mult, pres_fac_max=pmax, hist_cost_weight=hist_w,
        hist_gain=hist_gain, history_decay=1.0, hotset_cap=hotset_cap, hotset_percentage=hp,
        layer_bias_alpha=lb[0], layer_bias_min=lb[1], layer_bias_max=lb[2], via_cost_base=via_cost,
        stagnation_patience=patience, max_iterations=max_it, strategy=strat)


def apply_derived_parameters(config, derived: DerivedRoutingParameters):
    config.pres_fac_init = derived.pres_fac_init
    config.pres_fac_mult = derived.pres_fac_mult
    config.pres_fac_max = derived.pres_fac_max
    config.hist_cost_weight = derived.hist_cost_weight
    config.hist_gain = derived.hist_gain
    config.history_decay = derived.history_decay
    config.hotset_cap = derived.hotset_cap
    config.via_cost = derived.via_cost_base
    config.stagnation_patience = derived.stagnation_patience
    config.max_iterations = max(config.max_iterations, derived.max_iterations)
    cap = getattr(config, "max_iterations_cap", None)
    if cap is not None:
        config.max_iterations = int(cap)
