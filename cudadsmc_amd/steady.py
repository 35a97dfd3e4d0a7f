"""Steady-state detection for DSMC runs: the sliding-window change of the
sampled fields measured against the sampling noise the run itself shows.

A run is sampled in windows of W steps, each as two halves of W/2 steps
(the moment sums are read at the middle and at the end of a window).  Every
watched field q -- number density, flow velocity (both components together),
temperature -- gives per window w, with half-window values a_w and b_w per
cell, the window value m_w = (a_w + b_w)/2 and cell weights g (mostly-fluid
cells sampled in both halves):

    s_w^2 = <(b_w - a_w)^2>_g        noise (plus drift inside the window)
    d_w^2 = <(m_w - m_{w-k})^2>_g    change over the last k windows
    chi_q = sqrt(2 d_w^2 / ((s_w^2 + s_{w-k}^2) / 2))

and the window's chi is the largest over the fields.  Velocity and
temperature are weighted by the cell's sample count (their noise variance
falls as 1 / samples, so the sparse wake cells do not drown the rest);
density is unweighted.  The same weights enter the noise and the change, so
a steady flow gives chi near 1 for every field.

Measuring the noise from the two halves keeps the statistic right when
consecutive steps' samples are correlated (a particle stays in a cell for
several steps), which a Poisson count model misses.  When a drift r per
step dominates the noise, d ~ k rW and s ~ rW/2, so chi approaches
2 sqrt(2) k.  The lag k makes slow drifts -- smaller than the noise from one
window to the next, but adding up -- count.  Until k windows lie behind a
window it is compared with the first one (lag w - 1 at window w), so a run
that is steady from its start is confirmed after patience + 1 windows; from
window k + 1 on every chi uses the full lag.  The detector declares a run
steady at the first window that completes `patience` consecutive windows
with chi below `threshold`.
"""
import numpy as np

FIELDS = ("n", "u", "T")


def _as_fields(a):
    """One window-half as a list of per-field arrays (C,) or (k, C)."""
    if isinstance(a, np.ndarray) and a.ndim == 1:
        return [a.astype(np.float64)]
    return [np.asarray(f, np.float64) for f in a]


def _msq(x, w):
    """Weighted mean over cells of the squared (vector) difference x."""
    x2 = x * x if x.ndim == 1 else np.sum(x * x, 0)
    return float(np.sum(w * x2) / np.sum(w))


class SteadyDetector:
    def __init__(self, threshold=1.5, patience=3, lag=1):
        self.threshold = threshold
        self.patience = patience
        self.lag = lag
        self.past = []              # (field means, their noise, weights, mask) of the last windows
        self.below = 0
        self.windows = 0
        self.history = []           # chi of every window (largest over the fields)
        self.field_history = []     # chi of every field, per window
        self.steady_at = None       # window index (1-based, counting every window fed)

    def update(self, a, b, mask=None, weights=None):
        """Feeds one window: a and b are the two halves, either one density
        array per cell or a list of fields, each (C,) or (k, C) for a vector
        field; weights (one per field, None = uniform) weigh the cells.
        Returns True once steady."""
        fa, fb = _as_fields(a), _as_fields(b)
        nf = len(fa)
        C = fa[0].shape[-1]
        ok = np.ones(C, bool) if mask is None else np.asarray(mask, bool).ravel()
        wts = [None] * nf if weights is None else list(weights)
        m = [0.5 * (x + y) for x, y in zip(fa, fb)]
        w = [np.where(ok, 1.0 if g is None else np.asarray(g, np.float64), 0.0) for g in wts]
        s2 = [_msq((y - x)[..., ok], wi[ok]) if wi[ok].sum() > 0 else 0.0 for x, y, wi in zip(fa, fb, w)]
        self.windows += 1
        if self.past:
            pm, ps2, pw, pok = self.past[-min(self.lag, len(self.past))]
            both = ok & pok
            chis = []
            for q in range(nf):
                wq = np.minimum(w[q], pw[q])[both]
                if wq.sum() <= 0:
                    chis.append(float("inf"))
                    continue
                d2 = _msq((m[q] - pm[q])[..., both], wq)
                noise = 0.5 * (s2[q] + ps2[q])
                chis.append(float(np.sqrt(2.0 * d2 / noise)) if noise > 0 else float("inf"))
            chi = max(chis)
            self.history.append(chi)
            self.field_history.append(chis)
            self.below = self.below + 1 if chi < self.threshold else 0
            if self.steady_at is None and self.below >= self.patience:
                self.steady_at = self.windows
        self.past.append((m, s2, w, ok))
        self.past = self.past[-self.lag:]
        return self.steady_at is not None


def density(acc_counts, nsteps, Fn, vol):
    """Per-cell mean number density over nsteps steps from the sample count
    sums (particle-samples per cell)."""
    N = np.asarray(acc_counts, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(vol > 0, N / max(nsteps, 1) * Fn / np.where(vol > 0, vol, 1.0), 0.0)


def half_fields(sums, nsteps, Fn, vol, mass, boltz):
    """Density, flow velocity (ux, uy) and temperature per cell of one
    half-window from the differences of its moment sums (rows: sum 1, sum
    vx, sum vy, sum vz, sum v^2), and the sample count per cell.  Cells
    without samples get zeros."""
    N = np.asarray(sums[0], np.float64)
    has = N > 0
    Ns = np.where(has, N, 1.0)
    ux, uy, uz = (np.where(has, sums[k] / Ns, 0.0) for k in (1, 2, 3))
    T = np.where(has, mass / (3.0 * boltz) * (sums[4] / Ns - (ux * ux + uy * uy + uz * uz)), 0.0)
    return [density(N, nsteps, Fn, vol), np.stack([ux, uy]), T], N
