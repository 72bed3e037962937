from fftcorr_amd.catalog.synthetic import thomas_catalog, uniform_catalog
from fftcorr_amd.catalog.binary import read_catalog, read_header, write_catalog
from fftcorr_amd.catalog.random_catalog import add_random_particles
