"""Planted-phase inputs whose calibration the mathematics fixes.

The reference's inversion (algorithm/cpu/CpuLeapCalibrator.cc:159-178) turns
pol 0 of avgData into a calibration through

    phi1  = arg(avgData(I1, 0)) ++ 0                      (:161-163)
    cal1  = Ad1 * phi1                                    (:165)
    dInt  = arg(exp(-2 pi i (A cal1)(n)) avgData(n, 0))   (:168-172, row n, not row I(n))
    cal   = Ad * (dInt ++ 0) + cal1                       (:174-178)

When avgData(b, 0) = amp_b exp(i (phi[a1(b)] - phi[a2(b)])) for a per-antenna
phase screen phi, each stage has an answer that does not go through that
formula:

* A1 is the star around the reference antenna r plus the row e_r, so it is
  exactly solvable and cal1 = wrap(phi - phi[r]) (exactly phi - phi[r] when
  no baseline phase wraps; 0 for an antenna whose baseline to r is flagged).
* With every baseline unflagged and no autocorrelations, I(n) = n and
  (A cal1)(n) = phi[a1(n)] - phi[a2(n)], so dInt(n) = (1 - 2 pi)(A cal1)(n)
  as long as (2 pi - 1)|phi[a1] - phi[a2]| < pi.
* A's constraint row is e_0 (MetaData.cc:97 passes no reference antenna,
  PhaseMatrixFunction.cc:81-87), so A 1 = e_last and, for a connected array,
  Ad e_last = 1 and Ad [A_top x; 0] = x - x[0] 1.

Together: cal = (2 - 2 pi)(phi - phi[r]) - (1 - 2 pi)(phi[0] - phi[r]).
"""
import numpy as np

TWO_PI = 2.0 * np.pi


def wrap(x):
    return (np.asarray(x) + np.pi) % TWO_PI - np.pi


def baselines(n_ant, autos):
    """Antenna pairs in measurement-set order: (i, j) for j >= i (autos) or j > i."""
    a1, a2 = [], []
    for i in range(n_ant):
        for j in range(i if autos else i + 1, n_ant):
            a1.append(i)
            a2.append(j)
    return np.array(a1, dtype=np.int32), np.array(a2, dtype=np.int32)


def inversion_ms(n_ant, autos, flagged=(), n_pols=4):
    """A one-timestep, one-channel measurement set carrying only what MetaData
    reads: antenna pairs, first-epoch flags, counts.  The visibilities are
    zero (the inversion tests feed avgData directly)."""
    a1, a2 = baselines(n_ant, autos)
    nbl = len(a1)
    flag = np.zeros(nbl, dtype=np.uint8)
    flag[list(flagged)] = 1
    return {
        "num_stations": n_ant, "total_antennas": n_ant, "num_baselines": nbl, "num_rows": nbl,
        "num_channels": 1, "num_pols": n_pols, "freq_start_hz": 1.0e8, "freq_inc_hz": 1.0e6,
        "phase_centre": (0.3, -0.4), "antenna1": a1, "antenna2": a2, "time": np.zeros(nbl),
        "flag": flag, "uvw": np.zeros((nbl, 3)), "vis": np.zeros((nbl, 1, n_pols), dtype=np.complex128),
    }


def planted_avg(ms, phi, rng):
    """avgData [baseline][pol]: pol 0 carries the planted baseline phase with a
    random amplitude; the other pols are random (only pol 0 reaches cal)."""
    a1, a2 = ms["antenna1"], ms["antenna2"]
    nbl, P = ms["num_baselines"], ms["num_pols"]
    avg = rng.standard_normal((nbl, P)) + 1j * rng.standard_normal((nbl, P))
    avg[:, 0] = rng.uniform(0.5, 3.0, nbl) * np.exp(1j * (phi[a1] - phi[a2]))
    return avg


def closed_form_cal(phi, ref_ant):
    """cal for a connected array with no autocorrelations, no flags and baseline
    phases small enough that nothing wraps (module docstring)."""
    cal1 = phi - phi[ref_ant]
    return (2.0 - TWO_PI) * cal1 - (1.0 - TWO_PI) * cal1[0]


def star_cal1(ms, phi, ref_ant):
    """cal1 of the planted screen: the star around ref_ant is exactly solvable,
    so antenna j gets wrap(phi[j] - phi[r]) when its baseline to r is unflagged
    and not an autocorrelation, and 0 otherwise (a zero column of A1)."""
    a1, a2, flag = ms["antenna1"], ms["antenna2"], ms["flag"]
    cal1 = np.zeros(len(phi))
    for b in range(ms["num_baselines"]):
        if flag[b] or a1[b] == a2[b] or ref_ant not in (a1[b], a2[b]):
            continue
        j = a2[b] if a1[b] == ref_ant else a1[b]
        cal1[j] = wrap(phi[j] - phi[ref_ant])
    return cal1


def planted_cal(ms, phi, ref_ant, A, I, delta_rows=None, two_pi=TWO_PI, cal1_sign=1.0):
    """cal from the planted screen for any flags / autocorrelations:
    cal1 from the star solution, dInt(n) from the planted phase of avgData row n
    and row n of A, then the least-squares solve by LAPACK's pseudo-inverse.
    The keyword arguments model plausible misreadings: `delta_rows` picks which
    avgData row feeds dInt(n), `two_pi` the factor on A cal1, `cal1_sign` the
    sign with which cal1 is added back."""
    a1, a2 = ms["antenna1"], ms["antenna2"]
    cal1 = star_cal1(ms, phi, ref_ant)
    rows = np.arange(len(I)) if delta_rows is None else np.asarray(delta_rows)
    acal1 = cal1[a1[I]] - cal1[a2[I]]
    dint = wrap((phi[a1[rows]] - phi[a2[rows]]) - two_pi * acal1)
    return np.linalg.pinv(A) @ np.append(dint, 0.0) + cal1_sign * cal1


# ---------------------------------------------------------------------------
# A three-antenna set with autocorrelations, solved by hand
# ---------------------------------------------------------------------------
# Baselines in measurement-set order: 0 (0,0), 1 (0,1), 2 (0,2), 3 (1,1),
# 4 (1,2), 5 (2,2).  The default reference antenna is 2 (TotalAntennas - 1).
#
#   A1: rows (0,2), (1,2) and the constraint e_2, I1 = [2, 4]
#       -> cal1 = [p2, p4, 0] with p_b = arg avgData(b, 0)
#   A:  rows (0,1), (0,2), (1,2) and the constraint e_0, I = [1, 2, 4]
#       -> A cal1 = [p2 - p4, p2, p4]
#   dInt(n) = wrap(arg avgData(n, 0) - 2 pi (A cal1)(n)) for n = 0, 1, 2
#       (rows 0, 1, 2 of avgData, which are baselines (0,0), (0,1), (0,2),
#        while the rows of A are baselines (0,1), (0,2), (1,2))
#   least squares of A x = [d0, d1, d2, 0]: A^T A = [[3,-1,-1],[-1,2,-1],[-1,-1,2]],
#       (A^T A)^-1 = [[3,3,3],[3,5,4],[3,4,5]] / 3, A^T b = [d0 + d1, d2 - d0, -d1 - d2]
#       -> x = [0, (-2 d0 - d1 + d2) / 3, (-d0 - 2 d1 - d2) / 3]
#   cal = x + cal1
HAND_AVG = np.array([2.0, np.exp(0.3j), 1.5 * np.exp(0.1j), 1.0, np.exp(-0.2j), 0.5])


def hand_cal(avg0, rows=(0, 1, 2), two_pi=TWO_PI, cal1_sign=1.0):
    """cal of the three-antenna set by the hand solution above; `rows` are the
    avgData rows that feed dInt (the reference reads rows 0, 1, 2)."""
    p = np.angle(avg0)
    cal1 = np.array([p[2], p[4], 0.0])
    acal1 = np.array([cal1[0] - cal1[1], cal1[0] - cal1[2], cal1[1] - cal1[2]])
    d0, d1, d2 = (wrap(p[r] - two_pi * a) for r, a in zip(rows, acal1))
    x = np.array([0.0, (-2 * d0 - d1 + d2) / 3, (-d0 - 2 * d1 - d2) / 3])
    return x + cal1_sign * cal1
