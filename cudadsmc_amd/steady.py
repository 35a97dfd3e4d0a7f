"""Steady-state detection for DSMC runs: the sliding-window field change
measured against the sampling noise the run itself shows.

A run is sampled in windows of W steps, each as two halves of W/2 steps
(samples reset at the window start, read at the middle and at the end).
Per window w, with half-window mean densities a_w and b_w per cell and the
window mean m_w = (a_w + b_w)/2 (mostly-fluid, sampled cells):

    s_w^2 = mean_c (b_w - a_w)^2        noise (plus drift inside the window)
    d_w^2 = mean_c (m_w - m_{w-k})^2    change over the last k windows
    chi_w = sqrt(2 d_w^2 / ((s_w^2 + s_{w-k}^2) / 2))

Measuring the noise from the two halves keeps the statistic right when
consecutive steps' samples are correlated (a particle stays in a cell for
several steps), which a Poisson count model misses.  For a steady flow
chi is about 1; when a drift r per step dominates the noise, d ~ k rW and
s ~ rW/2, so chi approaches 2 sqrt(2) k.  The lag k makes slow drifts --
smaller than the noise from one window to the next, but adding up -- count.
Until k windows lie behind a window, it is compared with the first one
(lag w - 1 at window w), so a run that is steady from its start is
confirmed after patience + 1 windows rather than k + patience; from window
k + 1 on every chi uses the full lag.
The detector declares a run steady at the first window that completes
`patience` consecutive windows with chi below `threshold`.
"""
import numpy as np


class SteadyDetector:
    def __init__(self, threshold=1.5, patience=3, lag=1):
        self.threshold = threshold
        self.patience = patience
        self.lag = lag
        self.past = []              # (window mean, half-difference mean square, mask) of the last windows
        self.below = 0
        self.windows = 0
        self.history = []
        self.steady_at = None       # window index (1-based, counting every window fed)

    def update(self, a, b, mask=None):
        """Feeds one window's two half-window mean densities per cell;
        returns True once steady."""
        a = np.asarray(a, np.float64).ravel()
        b = np.asarray(b, np.float64).ravel()
        ok = np.ones(a.shape, bool) if mask is None else np.asarray(mask, bool).ravel()
        m = 0.5 * (a + b)
        s2 = float(np.mean((b[ok] - a[ok]) ** 2)) if ok.any() else 0.0
        self.windows += 1
        if self.past:
            pm, ps2, pok = self.past[-min(self.lag, len(self.past))]
            both = ok & pok
            d2 = float(np.mean((m[both] - pm[both]) ** 2)) if both.any() else float("inf")
            noise = 0.5 * (s2 + ps2)
            chi = float(np.sqrt(2.0 * d2 / noise)) if noise > 0 else float("inf")
            self.history.append(chi)
            self.below = self.below + 1 if chi < self.threshold else 0
            if self.steady_at is None and self.below >= self.patience:
                self.steady_at = self.windows
        self.past.append((m, s2, ok))
        self.past = self.past[-self.lag:]
        return self.steady_at is not None


def density(acc_counts, nsteps, Fn, vol):
    """Per-cell mean number density over nsteps steps from the sample count
    sums (particle-samples per cell)."""
    N = np.asarray(acc_counts, np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.where(vol > 0, N / max(nsteps, 1) * Fn / np.where(vol > 0, vol, 1.0), 0.0)
