"""Independent numpy restatements used to pin the oracle on small inputs.

These are written against the reference's documented behaviour, not against
the oracle's code: vectorised numpy, LAPACK (numpy.linalg.pinv) for the
pseudo-inverse, and closed forms where the mathematics gives one.
"""
import numpy as np

C_LIGHT = 299792458.0


def phase_matrix(a1, a2, fg, ref_ant=None):
    """PhaseMatrixFunction by brute force: rows for unflagged cross baselines
    (touching ref_ant when given), then a constraint row on ref_ant (or 0)."""
    a1 = np.asarray(a1)
    a2 = np.asarray(a2)
    fg = np.asarray(fg).astype(bool)
    n_ant = int(max(a1.max(), a2.max())) + 1
    keep = (a1 != a2) & ~fg
    if ref_ant is not None:
        keep &= (a1 == ref_ant) | (a2 == ref_ant)
    I = np.nonzero(keep)[0].astype(np.int32)
    A = np.zeros((len(I) + 1, n_ant))
    A[np.arange(len(I)), a1[I]] = 1.0
    A[np.arange(len(I)), a2[I]] = -1.0
    A[-1, 0 if ref_ant is None else ref_ant] = 1.0
    return A, I


def rotation_x(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[1, 0, 0], [0, c, -s], [0, s, c]])


def rotation_z_t(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, s, 0], [-s, c, 0], [0, 0, 1]])


def dd_matrix(centre, direction):
    return rotation_x(-np.pi / 2 + direction[1]) @ rotation_z_t(direction[0] - centre[0]) @ rotation_x(
        np.pi / 2 - centre[1])


def rotate_visibilities(ms, direction, row_start=0, n_rows=None):
    """avgData[bl][pol]: sum over the integration rows and channels of
    vis * exp(i shift), skipping (row, chan) with a NaN in any pol."""
    if n_rows is None:
        n_rows = ms["num_rows"] - row_start
    nbl = ms["num_baselines"]
    C = ms["num_channels"]
    vis = ms["vis"][row_start:row_start + n_rows].astype(np.complex128)
    uvw = ms["uvw"][row_start:row_start + n_rows]
    dd = dd_matrix(ms["phase_centre"], direction)
    wrot = uvw @ dd[2]
    shift_factor = -2 * np.pi * (wrot - uvw[:, 2])
    lam = C_LIGHT / (ms["freq_start_hz"] + np.arange(C) * ms["freq_inc_hz"])
    rot = vis * np.exp(1j * (shift_factor[:, None] / lam[None, :]))[:, :, None]
    ok = ~np.isnan(rot.real).any(axis=2) & ~np.isnan(rot.imag).any(axis=2)
    rot = np.where(ok[:, :, None], rot, 0)
    per_row = rot.sum(axis=1)  # [rows][pol]
    out = np.zeros((nbl, ms["num_pols"]), dtype=np.complex128)
    md = np.arange(n_rows) % nbl
    np.add.at(out, md, per_row)
    return out


def apply_inversion(avg, A, I, A1, I1, Ad=None, Ad1=None):
    """Inversion of PhaseRotate with LAPACK pseudo-inverses unless given."""
    Ad = np.linalg.pinv(A) if Ad is None else Ad
    Ad1 = np.linalg.pinv(A1) if Ad1 is None else Ad1
    nbl = avg.shape[0]
    idx = np.where(I1 < 0, I1 + nbl, I1)
    phi1 = np.append(np.angle(avg[idx, 0]), 0.0)
    cal1 = Ad1 @ phi1
    acal1 = A @ cal1
    n = len(I)
    dint = np.angle(np.exp(-2j * np.pi * acal1[:n]) * avg[:n, 0])
    dphi = np.append(dint, 0.0)
    return Ad @ dphi + cal1


def star_cal1(phi1, a1, a2, I1, ref_ant, n_ant):
    """Closed form of Ad1 * phi1: the star graph around ref_ant plus the
    constraint x[ref] = 0 is exactly solvable, x[j] = -phi for (ref, j) and
    x[j] = phi for (j, ref); antennas without a baseline to ref get 0."""
    x = np.zeros(n_ant)
    for k, b in enumerate(I1):
        if a1[b] == ref_ant:
            x[a2[b]] = -phi1[k]
        else:
            x[a1[b]] = phi1[k]
    return x


def wrap(x):
    return (np.asarray(x) + np.pi) % (2 * np.pi) - np.pi
