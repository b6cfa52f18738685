/*
 * vtk_writer.c -- VTK legacy/XML writers and the PVD time-series index for
 * the spectral fields, in the formats of the reference's
 * src/output/vtk_writer.c (see include/vtk_writer.h).
 */
#include "vtk_writer.h"

/* legacy ASCII structured grid (vtk_writer.c:10-58) */
int ns_vtk_write_spectral_legacy(const char *filename, const NSSpectralData *s, double time) {
    if (!filename || !s || !s->x || !s->y) return -1;
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    const int nx = s->nx, ny = s->ny;
    const long n = (long)nx * ny;
    fprintf(fp, "# vtk DataFile Version 3.0\nNavier-Stokes Solution t=%.6f\nASCII\n", time);
    fprintf(fp, "DATASET STRUCTURED_GRID\nDIMENSIONS %d %d 1\nPOINTS %ld double\n", nx, ny, n);
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) fprintf(fp, "%.9f %.9f 0.0\n", s->x[i], s->y[j]);
    fprintf(fp, "FIELD FieldData 1\nTIME 1 1 double\n%.9f\n", time);
    fprintf(fp, "POINT_DATA %ld\n", n);
    if (s->u && s->v) {
        fprintf(fp, "VECTORS velocity double\n");
        for (long q = 0; q < n; ++q) fprintf(fp, "%.9f %.9f 0.0\n", s->u[q], s->v[q]);
    }
    if (s->omega) {
        fprintf(fp, "SCALARS vorticity double 1\nLOOKUP_TABLE default\n");
        for (long q = 0; q < n; ++q) fprintf(fp, "%.9f\n", s->omega[q]);
    }
    return fclose(fp) == 0 ? 0 : -1;
}

/* XML rectilinear grid (vtk_writer.c:60-135) */
int ns_vtk_write_spectral_xml(const char *filename, const NSSpectralData *s, double time) {
    if (!filename || !s || !s->x || !s->y) return -1;
    FILE *fp = fopen(filename, "w");
    if (!fp) return -1;
    const int nx = s->nx, ny = s->ny;
    const int have_w = s->omega != NULL, have_uv = s->u && s->v;
    fprintf(fp, "<?xml version=\"1.0\"?>\n");
    fprintf(fp, "<VTKFile type=\"RectilinearGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
    fprintf(fp, "  <RectilinearGrid WholeExtent=\"0 %d 0 %d 0 0\">\n", nx - 1, ny - 1);
    fprintf(fp, "    <FieldData>\n");
    fprintf(fp, "      <DataArray type=\"Float64\" Name=\"TimeValue\" NumberOfTuples=\"1\" "
                "format=\"ascii\">%.9f</DataArray>\n", time);
    fprintf(fp, "    </FieldData>\n");
    fprintf(fp, "    <Piece Extent=\"0 %d 0 %d 0 0\">\n", nx - 1, ny - 1);
    fprintf(fp, "      <PointData%s%s>\n", have_w ? " Scalars=\"vorticity\"" : "",
            have_uv ? " Vectors=\"velocity\"" : "");
    if (have_w) {
        fprintf(fp, "        <DataArray type=\"Float64\" Name=\"vorticity\" format=\"ascii\">\n");
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) fprintf(fp, " %.9f", s->omega[(long)j * nx + i]);
            fprintf(fp, "\n");
        }
        fprintf(fp, "        </DataArray>\n");
    }
    if (have_uv) {
        fprintf(fp, "        <DataArray type=\"Float64\" NumberOfComponents=\"3\" Name=\"velocity\" "
                    "format=\"ascii\">\n");
        for (int j = 0; j < ny; ++j) {
            for (int i = 0; i < nx; ++i) {
                const long q = (long)j * nx + i;
                fprintf(fp, " %.9f %.9f 0.0", s->u[q], s->v[q]);
            }
            fprintf(fp, "\n");
        }
        fprintf(fp, "        </DataArray>\n");
    }
    fprintf(fp, "      </PointData>\n      <CellData/>\n      <Coordinates>\n");
    fprintf(fp, "        <DataArray type=\"Float64\" Name=\"X\" format=\"ascii\">\n");
    for (int i = 0; i < nx; ++i) fprintf(fp, " %.9f", s->x[i]);
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "        <DataArray type=\"Float64\" Name=\"Y\" format=\"ascii\">\n");
    for (int j = 0; j < ny; ++j) fprintf(fp, " %.9f", s->y[j]);
    fprintf(fp, "\n        </DataArray>\n");
    fprintf(fp, "        <DataArray type=\"Float64\" Name=\"Z\" format=\"ascii\"> 0.0 </DataArray>\n");
    fprintf(fp, "      </Coordinates>\n    </Piece>\n  </RectilinearGrid>\n</VTKFile>\n");
    return fclose(fp) == 0 ? 0 : -1;
}

/* PVD collection: create it, or insert one DataSet line before </Collection>
 * (vtk_writer.c:273-330) */
int ns_vtk_update_pvd(const char *pvd, const char *dataset_file, double time) {
    if (!pvd || !dataset_file) return -1;
    FILE *in = fopen(pvd, "r");
    if (!in) {
        FILE *fp = fopen(pvd, "w");
        if (!fp) return -1;
        fprintf(fp, "<?xml version=\"1.0\"?>\n"
                    "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
                    "  <Collection>\n"
                    "    <DataSet timestep=\"%.9f\" group=\"\" part=\"0\" file=\"%s\"/>\n"
                    "  </Collection>\n"
                    "</VTKFile>\n", time, dataset_file);
        return fclose(fp) == 0 ? 0 : -1;
    }
    size_t cap = 4096, len = 0;
    char *buf = (char *)malloc(cap);
    int ch;
    while (buf && (ch = fgetc(in)) != EOF) {
        if (len + 1 >= cap) {
            char *nb = (char *)realloc(buf, cap *= 2);
            if (!nb) {
                free(buf);
                buf = NULL;
                break;
            }
            buf = nb;
        }
        buf[len++] = (char)ch;
    }
    fclose(in);
    if (!buf) return -1;
    buf[len] = '\0';
    char *end = strstr(buf, "</Collection>");
    if (!end) {
        free(buf);
        return -1;
    }
    FILE *fp = fopen(pvd, "w");
    if (!fp) {
        free(buf);
        return -1;
    }
    fwrite(buf, 1, (size_t)(end - buf), fp);
    fprintf(fp, "    <DataSet timestep=\"%.9f\" group=\"\" part=\"0\" file=\"%s\"/>\n", time,
            dataset_file);
    fputs(end, fp);
    free(buf);
    return fclose(fp) == 0 ? 0 : -1;
}
