from fftcorr_amd.catalog.synthetic import thomas_catalog, uniform_catalog
from fftcorr_amd.catalog.binary import read_catalog, read_header, write_catalog
