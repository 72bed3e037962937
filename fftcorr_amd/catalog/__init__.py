from fftcorr_amd.catalog.synthetic import thomas_catalog, uniform_catalog
