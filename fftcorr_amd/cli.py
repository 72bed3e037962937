"""Command-line driver with the reference CLI's options and text output
(cc/fftcorr.cc:175-460), running on the GPU:

    python -m fftcorr_amd.cli -n 256 -ell 2 -r 200 -dr 10 -in DD.dat -in2 RR.dat \\
        -d 1 -out corrDD.dat.out

Options: -n/-ngrid, -n3/-ngrid3, -ell/-maxell, -r/-sep, -dr/-dsep, -kmax/-k,
-dk, -cell/-c, -in/-i, -in2/-i2, -out/-o, -periodic/-p, -dist/-d, -wisdom/-wf,
plus -precision fp64|fp32|mixed (default fp64).  Distribution scheme 2 (the
D12 wavelet, the reference's default) needs the wavelet's sample table
registered first (fftcorr_amd.particle_mesh.set_wavelet_table); otherwise pass
-d 0 (NGP), 1 (TSC) or 3 (CIC).  A TSC run uses the TSC-aliased window correction
and shot-noise factor, like the reference; CIC uses the CIC-aliased window.
"""

import os
import sys
import time

import numpy as np

from fftcorr_amd import fftw
from fftcorr_amd.catalog.binary import read_catalog, read_header
from fftcorr_amd.correlate import Correlator, PeriodicCorrelator
from fftcorr_amd.grid import ConfigSpaceGrid
from fftcorr_amd.particle_mesh import MassAssignor

USAGE = """FFTCORR: Error in command-line
   -n <int> (or -ngrid): FFT linear grid size for a cubic box
   -n3 <int> <int> <int> (or -ngrid3): FFT linear grid sizes for rectangle
   -ell <int> (or -maxell): Multipole to compute.
   -r <float> (or -sep): Max separation.
   -dr <float> (or -dsep): Binning of separation.
   -kmax <float>: Maximum wavenumber for power spectrum.
   -dk <float>: Binning of wavenumber.
   -periodic (or -p): Configure for cubic periodic box.
   -dist <int> (or -d): Distribution scheme (0 NGP, 1 TSC, 3 CIC).
   -in <filename>:  Input file name
   -in2 <filename>: Second input file name
   -out <filename>: Output file name, default to stdout
   -wisdom <filename>:  plan cache file name
   -precision fp64|fp32|mixed
"""


def parse(argv):
    o = dict(sep=-123.0, dsep=10.0, kmax=0.03, dk=0.01, maxell=0, ngrid_cube=256,
             periodic=False, scheme=2, ngrid=None, cell=-123.0, infile="/tmp/corrRR.dat",
             infile2=None, outfile=None, wisdom="wisdom_fftcorr_hip", precision="fp64")
    i = 0
    try:
        while i < len(argv):
            a = argv[i]
            if a in ("-ngrid", "-n"):
                o["ngrid_cube"] = int(argv[i + 1]); i += 1
            elif a in ("-ngrid3", "-n3"):
                o["ngrid"] = [int(argv[i + 1]), int(argv[i + 2]), int(argv[i + 3])]; i += 3
            elif a in ("-maxell", "-ell"):
                o["maxell"] = int(argv[i + 1]); i += 1
            elif a in ("-sep", "-r"):
                o["sep"] = float(argv[i + 1]); i += 1
            elif a in ("-dsep", "-dr"):
                o["dsep"] = float(argv[i + 1]); i += 1
            elif a in ("-kmax", "-k"):
                o["kmax"] = float(argv[i + 1]); i += 1
            elif a == "-dk":
                o["dk"] = float(argv[i + 1]); i += 1
            elif a in ("-cell", "-c"):
                o["cell"] = float(argv[i + 1]); i += 1
            elif a in ("-in", "-i"):
                o["infile"] = argv[i + 1]; i += 1
            elif a in ("-in2", "-i2"):
                o["infile2"] = argv[i + 1]; i += 1
            elif a in ("-out", "-o"):
                o["outfile"] = argv[i + 1]; i += 1
            elif a in ("-periodic", "-p"):
                o["periodic"] = True
            elif a in ("-dist", "-d"):
                o["scheme"] = int(argv[i + 1]); i += 1
            elif a in ("-wisdom", "-wf"):
                o["wisdom"] = argv[i + 1]; i += 1
            elif a == "-precision":
                o["precision"] = argv[i + 1]; i += 1
            else:
                raise ValueError(a)
            i += 1
    except (IndexError, ValueError):
        sys.stderr.write(USAGE)
        raise SystemExit(1)
    if o["ngrid"] is None:
        o["ngrid"] = [o["ngrid_cube"]] * 3
    checks = [o["ngrid_cube"] > 0, o["maxell"] >= 0 and o["maxell"] % 2 == 0, o["sep"] > 0,
              o["dsep"] > 0, o["kmax"] != 0, o["dk"] > 0,
              not (o["periodic"] and o["cell"] >= 0), all(n > 0 for n in o["ngrid"])]
    if not all(checks):
        sys.stderr.write(USAGE)
        raise SystemExit(1)
    return o


def hist_lines(bins, counts, hist, prefix, normalize):
    """Rows as print_hist writes them (cc/fftcorr.cc:213-228)."""
    out = []
    for j in range(len(bins)):
        count = float(counts[0, j])
        denom = count if (normalize and count != 0) else 1.0
        line = "%1d %7.4f %8.0f" % (prefix, bins[j], count)
        line += "".join(" %16.9e" % (hist[i, j] / denom) for i in range(hist.shape[0]))
        out.append(line)
    return out


def grid_geometry(posmin, posmax, ngrid, cell, sep, periodic):
    """Grid corner and cell of the CLI (cc/fftcorr.cc:335-346): a
    non-periodic box is padded by sep / 1.5 on every side; the cell is the
    requested one or the smallest that covers the box, whichever is larger."""
    posmin = np.asarray(posmin, np.float64)
    posmax = np.asarray(posmax, np.float64)
    if not periodic:
        pad = sep / 1.5
        posmin = posmin - pad
        posmax = posmax + pad
    for i in range(3):
        cell = max(cell, (posmax[i] - posmin[i]) / ngrid[i])
    return posmin, cell


def run(argv, stdout=None):
    o = parse(argv)
    from fftcorr_amd.particle_mesh import wavelet_table_set
    if o["scheme"] not in (0, 1, 2, 3) or (o["scheme"] == 2 and not wavelet_table_set()):
        raise SystemExit("distribution scheme %d is not available on the GPU (use -d 0 NGP, "
                         "1 TSC or 3 CIC; -d 2, the D12 wavelet, needs its sample table "
                         "registered with fftcorr_amd.particle_mesh.set_wavelet_table)"
                         % o["scheme"])
    t0 = time.perf_counter()
    w = []
    w.append("# Running on the GPU (fftcorr_amd).")
    if os.path.exists(o["wisdom"]):
        fftw.import_wisdom_from_file(o["wisdom"])
    posmin, posmax = read_header(o["infile"])
    sep = o["sep"]
    ngrid = o["ngrid"]
    posmin, cell = grid_geometry(posmin, posmax, ngrid, o["cell"], sep, o["periodic"])
    w.append("# Adopted cell_size=%f for ngrid=[%d, %d, %d]" % (cell, *ngrid))
    w.append("# Adopted boxsize: %6.1f %6.1f %6.1f" % tuple(cell * n for n in ngrid))
    # The adopted cell is taken as given, as the reference's CLI builds its
    # grid from (ngrid, posmin, cell): recomputing it from posmax could round
    # one ulp above the adopted value.
    grid = ConfigSpaceGrid.from_cell(ngrid, posmin, cell, precision=o["precision"])
    ma = MassAssignor(grid, o["scheme"], o["periodic"], 1000000)
    for f in [o["infile"]] + ([o["infile2"]] if o["infile2"] else []):
        w.append("# Reading from file named %s" % f)
        _, _, rec = read_catalog(f)
        ma.add_particles(rec)
        w.append("# Found %d galaxies in this file" % rec.shape[0])
    w.append("# Found %d particles. Total weight %10.4e." % (ma.num_added, ma.totw))
    s1, s2 = grid.sum(), grid.sumsq()
    w.append("# Sum of grid is %10.4e (delta = %10.4e)" % (s1, s1 - ma.totw))
    w.append("# Sum of squares of grid is %10.4e " % s2)
    ncell = float(np.prod(ngrid))
    if o["periodic"]:
        mean = ma.totw / ncell
        w.append("# Normalizing by mean cell density %10.4e" % mean)
        grid.normalize(mean)
    totwsq = ma.totwsq
    sumsq = grid.sumsq()
    w.append("# Sum of squares of density = %14.7e" % sumsq)
    w.append("# Sum of squares of weights (divide by I for Pshot) = %14.7e" % totwsq)
    if o["scheme"] == 0:
        w.append("# Using nearest cell method")
    elif o["scheme"] == 1:
        totwsq *= 0.55 * 0.55 * 0.55
        w.append("# Using triangular cloud-in-cell")
    elif o["scheme"] == 2:
        w.append("# Using D12 wavelet")
    else:
        w.append("# Using cloud-in-cell")
    vcell = cell ** 3
    w.append("# Estimate of I (denominator) = %14.7e - %14.7e = %14.7e"
             % (sumsq / vcell, totwsq / vcell, (sumsq - totwsq) / vcell))
    window = {1: 2, 3: 4}.get(o["scheme"], 0)
    if o["periodic"]:
        corr = PeriodicCorrelator(ngrid, cell, window, sep, o["dsep"], o["kmax"], o["dk"],
                                  o["maxell"], precision=o["precision"])
    else:
        corr = Correlator(ngrid, cell, posmin, window, sep, o["dsep"], o["kmax"], o["dk"],
                          o["maxell"], precision=o["precision"])
    corr.autocorrelate(grid)
    r = corr.results()
    w.append("# Anisotropic power spectrum:")
    w += hist_lines(r["k"], r["k_counts"], r["k_hist"], 1, True)
    w.append("# Anisotropic correlations:")
    w += hist_lines(r["r"], r["r_counts"], r["r_hist"], 0, o["periodic"])
    w.append("#\n# Zero-lag correlations are %14.7e" % r["zerolag"])
    sum_ell0 = float(np.sum(r["k_hist"][0]))
    w.append("#\n# Integral of power spectrum is %14.7e" % (sum_ell0 / (cell ** 3 * ncell)))
    fftw.export_wisdom_to_file(o["wisdom"])
    w.append("#")
    sys.stderr.write("Total time %.3f s (GPU)\n" % (time.perf_counter() - t0))
    text = "\n".join(w) + "\n"
    if o["outfile"]:
        with open(o["outfile"], "w") as f:
            f.write(text)
    else:
        (stdout or sys.stdout).write(text)
    return text


if __name__ == "__main__":
    run(sys.argv[1:])
