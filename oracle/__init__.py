"""CPU oracle for the fftcorr correlation path.  TEST INFRASTRUCTURE ONLY:
imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
never by the fftcorr_amd package."""
