#!/usr/bin/env python3
"""Converts a casacore MeasurementSet into a raw measurement-set directory
(the format icrar/leap-accelerate/ms/RawMeasurementSet.h reads), for
machines where casacore is installed (python-casacore).

It reads exactly what the reference's calibration path reads:
  main table   ANTENNA1, ANTENNA2, TIME, UVW (ms/MeasurementSet.cc:105-130,
               287-294, MetaData.cc:86-89), FLAG[row][channel 0][correlation 0]
               (MeasurementSet.cc:181-197) and the visibility column
               (DATA, CORRECTED_DATA or MODEL_DATA; ms/utils.h:81-146),
  ANTENNA      its row count (MeasurementSet.cc:110-113),
  SPECTRAL_WINDOW  REF_FREQUENCY[0], CHAN_WIDTH[0][0], NUM_CHAN[0]
               (MetaData.cc:58-60, MeasurementSet.cc:169-179),
  POLARIZATION NUM_CORR[0] (MeasurementSet.cc:137-147),
  FIELD        PHASE_DIR[0] (MetaData.cc:67-79).
The visibility column is copied in chunks of rows in its own type (complex64
for a standard MS), so a cube larger than memory converts.

    python3 ms_to_raw.py <input.ms> <output-dir> [--column DATA] [--chunk-rows N]
"""
import argparse
import json
import os
import sys

import numpy as np


def convert(ms_path, out_dir, column="DATA", chunk_rows=65536):
    from casacore import tables  # python-casacore

    if column not in ("DATA", "CORRECTED_DATA", "MODEL_DATA"):
        raise ValueError("ms column not found")  # ms/utils.h:98-109
    main = tables.table(ms_path, readonly=True, ack=False)
    try:
        rows = main.nrows()
        ant = tables.table(os.path.join(ms_path, "ANTENNA"), readonly=True, ack=False)
        spw = tables.table(os.path.join(ms_path, "SPECTRAL_WINDOW"), readonly=True, ack=False)
        pol = tables.table(os.path.join(ms_path, "POLARIZATION"), readonly=True, ack=False)
        field = tables.table(os.path.join(ms_path, "FIELD"), readonly=True, ack=False)
        try:
            antenna_rows = ant.nrows()
            nchan = int(spw.getcol("NUM_CHAN")[0]) if spw.nrows() > 0 else 0
            ref_freq = float(spw.getcol("REF_FREQUENCY")[0])
            chan_width = float(np.asarray(spw.getcol("CHAN_WIDTH")[0]).ravel()[0])
            if pol.nrows() == 0:
                raise NotImplementedError("measurement set has no POLARIZATION rows")  # MeasurementSet.cc:143-146
            ncorr = int(pol.getcol("NUM_CORR")[0])
            phase_dir = (0.0, 0.0)
            if field.nrows() > 0:
                d = np.asarray(field.getcol("PHASE_DIR")[0]).reshape(-1, 2)[0]
                phase_dir = (float(d[0]), float(d[1]))
        finally:
            for t in (ant, spw, pol, field):
                t.close()

        os.makedirs(out_dir, exist_ok=True)
        np.asarray(main.getcol("ANTENNA1"), dtype="<i4").tofile(os.path.join(out_dir, "ANTENNA1.i32"))
        np.asarray(main.getcol("ANTENNA2"), dtype="<i4").tofile(os.path.join(out_dir, "ANTENNA2.i32"))
        np.asarray(main.getcol("TIME"), dtype="<f8").tofile(os.path.join(out_dir, "TIME.f64"))
        np.asarray(main.getcol("UVW"), dtype="<f8").reshape(rows, 3).tofile(os.path.join(out_dir, "UVW.f64"))
        dtype = None
        with open(os.path.join(out_dir, "FLAG_C0P0.u8"), "wb") as ff, open(os.path.join(out_dir, "DATA.part"), "wb") as fd:
            for r0 in range(0, rows, chunk_rows):
                n = min(chunk_rows, rows - r0)
                flag = np.asarray(main.getcol("FLAG", startrow=r0, nrow=n))  # [row][channel][correlation]
                np.ascontiguousarray(flag[:, 0, 0], dtype=np.uint8).tofile(ff)
                data = np.asarray(main.getcol(column, startrow=r0, nrow=n))
                if data.ndim != 3 or data.shape[1] != nchan or data.shape[2] != ncorr:
                    raise ValueError(f"{column} rows have shape {data.shape[1:]}, expected ({nchan}, {ncorr})")
                if dtype is None:
                    dtype = np.complex128 if data.dtype == np.complex128 else np.complex64
                np.ascontiguousarray(data, dtype=dtype).tofile(fd)
        dtype = dtype or np.complex64
        os.replace(os.path.join(out_dir, "DATA.part"),
                   os.path.join(out_dir, "DATA.c64" if dtype == np.complex64 else "DATA.c128"))
        header = {"format": "leap-raw-ms", "version": 1, "rows": int(rows), "channels": nchan, "pols": ncorr,
                  "antenna_table_rows": int(antenna_rows), "ref_frequency_hz": ref_freq, "chan_width_hz": chan_width,
                  "phase_dir_rad": [phase_dir[0], phase_dir[1]], "data_column": column,
                  "data_type": "complex64" if dtype == np.complex64 else "complex128"}
        with open(os.path.join(out_dir, "header.json"), "w") as f:
            json.dump(header, f)
        return header
    finally:
        main.close()


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("ms")
    ap.add_argument("out")
    ap.add_argument("--column", default="DATA")
    ap.add_argument("--chunk-rows", type=int, default=65536)
    a = ap.parse_args(argv)
    h = convert(a.ms, a.out, a.column, a.chunk_rows)
    print(json.dumps(h))
    return 0


if __name__ == "__main__":
    sys.exit(main())
