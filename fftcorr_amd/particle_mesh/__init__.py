from fftcorr_amd.particle_mesh.mass_assignor import (MassAssignor, set_wavelet_table,
                                                      wavelet_table_set)
