/*
 * vtk_writer.h -- VTK output of the spectral solver's physical fields.
 *
 * The spectral half of the reference's src/output/vtk_writer.h:19-30 with
 * the same signatures and file formats (vtk_writer.c:10-58 legacy ASCII
 * STRUCTURED_GRID, :60-135 XML RectilinearGrid, :273-330 PVD collection):
 * points from data->x, data->y; vectors "velocity" = (u, v, 0); scalars
 * "vorticity" = omega; the time as field data.  Values are written with
 * %.9f as in the reference.  The finite-difference writers are out of scope.
 * The host arrays are written as they are (call ns_spectral_sync_host after
 * device-resident steps).  All return 0 on success, -1 on failure.
 */
#ifndef NS_VTK_WRITER_H
#define NS_VTK_WRITER_H

#include "ns_spectral_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

int ns_vtk_write_spectral_legacy(const char *filename, const NSSpectralData *state, double time);
int ns_vtk_write_spectral_xml(const char *filename, const NSSpectralData *state, double time);
int ns_vtk_update_pvd(const char *pvd_filename, const char *dataset_file, double time);

#ifdef __cplusplus
}
#endif

#endif /* NS_VTK_WRITER_H */
