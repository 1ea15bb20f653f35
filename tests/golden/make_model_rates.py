#!/usr/bin/env python3
"""Independent statistical fixture for the whole network model.

A float64 numpy implementation of the reference method (Crypto-TII/GPU4SNN
gpu4snn.cu: parameters cu:286-320, update cu:352-416, regimes cu:362-374,
fixed out-degree connectivity and weights cu:876-899, propagation into the
next step cu:323-331/420-467) driven by numpy's own PCG64 generator -- no
code shared with the oracle.  It records the firing rate of N=2500,
Nsyn=1000 networks for the three regimes over several seeds, and the rate of
the classic all-to-all network of Izhikevich (2003) as a second anchor.
Writes model_rates.json next to this file.
"""
import json
import os

import numpy as np


def simulate(n, nsyn, ge, gi, steps, seed):
    rng = np.random.default_rng(seed)
    ne = (4 * n) // 5
    r = rng.random(n)
    exc = np.arange(n) < ne
    a = np.where(exc, 0.02, 0.02 + 0.08 * r)
    b = np.where(exc, 0.2, 0.25 - 0.05 * r)
    c = np.where(exc, -65 + 15 * r**2, -65.0)
    d = np.where(exc, 8 - 6 * r**2, 2.0)
    v = -65 * np.ones(n)
    u = b * v
    post = np.empty((n, nsyn), np.int64)
    for i in range(n):
        cand = np.delete(np.arange(n), i)
        post[i] = rng.choice(cand, nsyn, replace=False)
    w = np.where(exc[:, None], 0.5 * rng.random((n, nsyn)), -rng.random((n, nsyn))) * (1000.0 / nsyn)
    g = np.where(exc, ge, gi)
    isyn = np.zeros(n)
    total = 0
    for _ in range(steps):
        I = g * rng.standard_normal(n) + isyn
        isyn = np.zeros(n)
        for _h in range(2):
            v = v + 0.5 * (0.04 * v * v + 5 * v + 140 - u + I)
        u = u + a * (b * v - u)
        f = (v >= 30) | ~np.isfinite(v)
        total += int(f.sum())
        v[f] = c[f]
        u[f] += d[f]
        idx = np.nonzero(f)[0]
        if idx.size:
            np.add.at(isyn, post[idx].ravel(), w[idx].ravel())
    return total / n / (steps / 1000.0)


def classic_izhikevich_2003(steps, seed):
    """Izhikevich (2003) 'Simple model of spiking neurons', Fig. 3 network:
    800 excitatory + 200 inhibitory neurons, all-to-all."""
    rng = np.random.default_rng(seed)
    ne, ni = 800, 200
    re, ri = rng.random(ne), rng.random(ni)
    a = np.r_[0.02 * np.ones(ne), 0.02 + 0.08 * ri]
    b = np.r_[0.2 * np.ones(ne), 0.25 - 0.05 * ri]
    c = np.r_[-65 + 15 * re**2, -65 * np.ones(ni)]
    d = np.r_[8 - 6 * re**2, 2 * np.ones(ni)]
    S = np.c_[0.5 * rng.random((ne + ni, ne)), -rng.random((ne + ni, ni))]
    v = -65 * np.ones(ne + ni)
    u = b * v
    total = 0
    for _ in range(steps):
        I = np.r_[5 * rng.standard_normal(ne), 2 * rng.standard_normal(ni)]
        fired = np.nonzero(v >= 30)[0]
        total += fired.size
        v[fired] = c[fired]
        u[fired] += d[fired]
        I = I + S[:, fired].sum(1)
        v = v + 0.5 * (0.04 * v**2 + 5 * v + 140 - u + I)
        v = v + 0.5 * (0.04 * v**2 + 5 * v + 140 - u + I)
        u = u + a * (b * v - u)
    return total / (ne + ni) / (steps / 1000.0)


def main():
    out = {"n": 2500, "nsyn": 1000, "steps": 2000, "seeds": [11, 12, 13], "rates_hz": {}}
    for name, (ge, gi) in {"quiet": (2.5, 1.0), "balanced": (5.0, 2.0), "irregular": (7.5, 3.0)}.items():
        out["rates_hz"][name] = [simulate(2500, 1000, ge, gi, 2000, s) for s in out["seeds"]]
        print(name, out["rates_hz"][name], flush=True)
    out["classic_2003_balanced_hz"] = [classic_izhikevich_2003(2000, s) for s in out["seeds"]]
    print("classic", out["classic_2003_balanced_hz"])
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "model_rates.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
