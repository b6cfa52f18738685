/*
 * checkpoint.h -- HDF5 checkpoint and restart of the spectral solver state.
 *
 * Replaces the spectral half of the reference's src/common/checkpoint.h:14-15
 * (ns_checkpoint_write_spectral / ns_checkpoint_read_spectral) with the same
 * signatures and the same file layout (src/common/checkpoint.c:266-302):
 * file attributes solver="spectral", time, dt, Re, nx, ny, Lx, Ly (written in
 * that order); group /fields with u, v, omega, psi as [ny][nx] doubles and
 * u_hat, v_hat, omega_hat as [ny][nx/2+1][2] doubles.
 *
 * Added so that a forced run resumes bit for bit: dataset psi_hat, dataset
 * sfd_hat when selective frequency damping is on, file attribute
 * forcing_calls (the counter of the forcing generator), dataset forcing_ou
 * (the coloured-noise state) when coloured forcing is on, and the attribute
 * cfl_factor.  The reader accepts files without these (reference files).
 *
 * HDF5 is loaded at run time (libhdf5.so, or the one under /opt/conda/lib);
 * without it both calls return -1 and ns_checkpoint_error() says so.  The
 * finite-difference checkpoint functions of the reference are out of scope.
 *
 * Write takes the host arrays as they are (call ns_spectral_sync_host after
 * device-resident steps); read fills the host arrays, t and dt, and sets the
 * device's forcing counter and OU state; ns_spectral_sync_device (or the next
 * ns_spectral_advance_rk4) uploads the state.  Both return 0 on success.
 */
#ifndef NS_CHECKPOINT_H
#define NS_CHECKPOINT_H

#include "ns_spectral_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

int ns_checkpoint_write_spectral(const char *filename, const NSSpectralData *state);
int ns_checkpoint_read_spectral(const char *filename, NSSpectralData *state);
const char *ns_checkpoint_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NS_CHECKPOINT_H */
