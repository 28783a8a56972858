"""Shared parity metrics: fitted parameters of two FIT_DTYPE arrays (TEST HELPERS)."""
import numpy as np

from localizer_amd import abi

NAMES = {4: ["A", "x0", "y0", "b"], 5: ["A", "r", "x0", "y0", "b"], 7: ["A", "sx", "sy", "x0", "y0", "rho", "b"]}
X_INDEX = {4: 1, 5: 2, 7: 3}


def compare(ref, got, model):
    """Returns a dict of agreement statistics between reference and candidate fits."""
    p = abi.n_params(model)
    xi = X_INDEX[p]
    acc_r = ref["status"] == 0
    acc_g = got["status"] == 0
    both = acc_r & acc_g
    P = ref["params"][both][:, :p]
    G = got["params"][both][:, :p]
    out = {
        "n": int(len(ref)),
        "status_agreement": float(np.mean(ref["status"] == got["status"])),
        "accept_agreement": float(np.mean(acc_r == acc_g)),
        "accepted_ref": int(acc_r.sum()),
        "accepted_got": int(acc_g.sum()),
        "both": int(both.sum()),
    }
    if both.sum() == 0:
        return out
    dxy = np.abs(P[:, xi:xi + 2] - G[:, xi:xi + 2]).max(axis=1)
    out["xy_max_px"] = float(dxy.max())
    out["xy_frac_within_1e-3"] = float(np.mean(dxy <= 1e-3))
    rel = np.abs(P - G) / np.maximum(np.abs(P), 1e-300)
    for i, nm in enumerate(NAMES[p]):
        out[f"{nm}_max_rel"] = float(rel[:, i].max())
        out[f"{nm}_frac_within_1e-4"] = float(np.mean(rel[:, i] <= 1e-4))
        out[f"{nm}_max_abs"] = float(np.abs(P[:, i] - G[:, i]).max())
    out["iterations_equal"] = float(np.mean(ref["iterations"][both] == got["iterations"][both]))
    return out


def format_stats(label, st):
    keys = sorted(k for k in st if k not in ("n",))
    return f"[{label}] n={st['n']} " + " ".join(f"{k}={st[k]:.6g}" if isinstance(st[k], float) else f"{k}={st[k]}"
                                              for k in keys)
