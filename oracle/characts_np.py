"""Trace characteristics of pyofss's Storage.update_characts
(pyofss/modules/storage.py:194-209), test infrastructure only.

Two forms of the same quantities:

* `reference_*`: the reference's computation itself, through the same
  scipy.signal routines it calls (field.py:207-260: find_peaks with height or
  prominence filters, peak_widths at rel_height 0.5, the find_x rule of
  spectrum_width_params).
* `restated_*`: the algorithm the device kernel implements, written with
  plain loops and reductions from scipy's documented definitions:
    - local maxima: samples whose two neighbours are smaller; a flat peak
      (equal samples) counts once, at the middle sample (rounded down); the
      first and last samples are never peaks;
    - prominence: from the peak, extend left and right while the signal is
      <= the peak height (equal heights do not stop the line), to the window
      border or a higher sample; on each side the minimum is the base (ties
      resolved towards the peak); prominence = height - max(bases);
    - width at rel_height 0.5: h = height - prominence / 2; from the peak
      move outwards while the sample exceeds h, not past the base; the end
      point is linearly interpolated where the sample is below h.
  tests/test_characts_oracle.py checks the restated form against the
  reference form on synthetic and propagated signals.
"""

import numpy as np
from scipy.signal import find_peaks, peak_widths


# ----------------------------------------------------------------------------
# reference (scipy) form, field.py:207-260
# ----------------------------------------------------------------------------
def reference_duration(P):
    """max_peak_params + get_duration (field.py:207-229, 283-287) in samples."""
    peaks, props = find_peaks(P, height=0)
    k = int(np.argmax(props["peak_heights"]))
    w = peak_widths(P, [peaks[k]], rel_height=0.5)
    return float(w[0][0])


def reference_spec_width(P):
    """spectrum_width_params + get_duration_spec (field.py:232-270, 276-280)
    in samples; None where the reference would fail (no qualifying peak)."""
    prominence = np.amax(P) / 100

    def find_x(y_peak, x_peak, is_left, ydata):
        x = np.where(ydata < y_peak)[0]
        filt = x[x < x_peak] if is_left else x[x > x_peak]
        if len(filt) == 0:
            return 0
        return int(np.amax(filt) if is_left else np.amin(filt))

    peaks, _ = find_peaks(P, prominence=prominence)
    if len(peaks) > 1:
        h = np.amax(P) / 10
        peaks.sort()
        return float(abs(find_x(h, peaks[0], True, P) - find_x(h, peaks[-1], False, P)))
    if len(peaks) == 0:
        return None
    return float(abs(peak_widths(P, peaks, rel_height=0.5)[0][0]))


def reference_peaks(P):
    """get_peaks (field.py:272-274)."""
    peaks, _ = find_peaks(P, height=0, prominence=(np.amax(P) / 10))
    return peaks


# ----------------------------------------------------------------------------
# restated form (the device algorithm)
# ----------------------------------------------------------------------------
def local_maxima(x):
    n = len(x)
    out = []
    i = 1
    while i < n - 1:
        if x[i - 1] < x[i]:
            j = i + 1
            while j < n - 1 and x[j] == x[i]:
                j += 1
            if x[j] < x[i]:
                out.append((i + j - 1) // 2)
            i = j
        else:
            i += 1
    return out


def prominence(x, p):
    """(prominence, left_base, right_base) of peak p."""
    h = x[p]
    i, lmin, lb = p, h, p
    while i >= 0 and x[i] <= h:
        if x[i] < lmin:
            lmin, lb = x[i], i
        i -= 1
    i, rmin, rb = p, h, p
    while i < len(x) and x[i] <= h:
        if x[i] < rmin:
            rmin, rb = x[i], i
        i += 1
    return h - max(lmin, rmin), lb, rb


def width(x, p, prom, lb, rb, rel=0.5):
    h = x[p] - prom * rel
    i = p
    while lb < i and h < x[i]:
        i -= 1
    left = float(i)
    if x[i] < h:
        left += (h - x[i]) / (x[i + 1] - x[i])
    i = p
    while i < rb and h < x[i]:
        i += 1
    right = float(i)
    if x[i] < h:
        right -= (h - x[i]) / (x[i - 1] - x[i])
    return right - left


def restated_duration(P):
    peaks = local_maxima(P)
    k = max(range(len(peaks)), key=lambda j: (P[peaks[j]], -j))
    p = peaks[k]
    prom, lb, rb = prominence(P, p)
    return width(P, p, prom, lb, rb)


def _prominent(P, thr):
    out = []
    floor = np.min(P)
    for p in local_maxima(P):
        if P[p] - floor < thr:  # prominence <= height - global minimum
            continue
        prom, lb, rb = prominence(P, p)
        if prom >= thr:
            out.append((p, prom, lb, rb))
    return out


def restated_spec_width(P):
    pk = _prominent(P, np.max(P) / 100)
    if len(pk) == 0:
        return None
    if len(pk) == 1:
        p, prom, lb, rb = pk[0]
        return abs(width(P, p, prom, lb, rb))
    h = np.max(P) / 10
    first, last = pk[0][0], pk[-1][0]
    below = np.nonzero(P < h)[0]
    lft = below[below < first]
    rgt = below[below > last]
    left = int(lft.max()) if len(lft) else 0
    right = int(rgt.min()) if len(rgt) else 0
    return float(abs(right - left))


def restated_peaks(P):
    return np.array([p for p, _, _, _ in _prominent(P, np.max(P) / 10)], dtype=np.int64)
