/* Writes the HDF5 fixtures of the field-file reader with the HDF5 library (1.10, as the reference
 * uses), in the reference's layout:
 *   - the fields: one dataset per field, named "u", "v", "w", "p" (and "t" with temperature),
 *     IEEE F64 big-endian, extents (2, my, mz, mx1+1) in Fortran order, i.e. C-order
 *     [mode][z][y][re/im] (h5dump_parallel, readwrite_hdf5.f90:650-722; save_4fields_fou_hdf5
 *     :266-381);
 *   - the header, written after the fields by reopening the file: 1-element I32LE and F64LE
 *     datasets, FORTRAN_S1 wall types and the bcy/bcz wall data (writeheader_hdf5, :853-918).
 *
 * Files (values are dyadic, so every element is exact in double and numpy recomputes them):
 *   uvwp_m17.h5      m = 17, 3 modes; default format (superblock 0, symbol-table group)
 *   uvwp_m25_poly.h5 m = 25, 2 modes; Chebyshev polynomials up to degree 19 plus T_20, T_21
 *                    and T_22 terms, the source of the interpolation tests; no "mx1" in
 *                    its header
 *   uvwpt_m9.h5      m = 9, 2 modes, temperature header; newest format (superblock 3, version-2
 *                    object headers, compact links) behind a 512-byte user block
 *
 * Build and run: tests/golden/make_uvwp_h5.sh (needs libhdf5; the files are committed). */
#include <hdf5.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static void check(herr_t e, const char* what)
{
    if (e < 0) {
        fprintf(stderr, "HDF5 call failed: %s\n", what);
        exit(1);
    }
}

/* Element [k][z][y][part] of field f in uvwp_m17.h5 and uvwpt_m9.h5. */
static double exact_value(int f, int k, int z, int y, int part)
{
    double v = (f + 1) * 1.0 + k * 0.5 + z * 0.03125 + y * 0.0009765625;
    if (part && k == 0) return 0.0;      /* the x-mean of a real field is real */
    return part ? -(v + 0.25) : v;
}

static double cheb(int n, double x) { return cos(n * acos(x)); }

/* The polynomial of uvwp_m25_poly.h5 at (y, z), mode k, part p, field f. */
static double poly_value(int f, int k, int part, double y, double z)
{
    double s = (f + 1) * (k + 1) * (part ? (k ? -0.5 : 0.0) : 1.0);   /* real x-mean */
    double v = 1.0 + 0.5 * cheb(1, y) - 0.25 * cheb(2, z) + 0.125 * cheb(3, y) * cheb(1, z)
             + 0.0625 * cheb(4, y) * cheb(4, z) + 0.03125 * cheb(19, y) * cheb(2, z)
             + 0.1 * cheb(20, z) + 0.3 * cheb(22, y) + 0.2 * cheb(1, y) * cheb(21, z);
    return s * v;
}

static void write_field(hid_t fid, const char* name, int my, int mz, int nm, const double* buf)
{
    hsize_t dims[4] = {(hsize_t)nm, (hsize_t)mz, (hsize_t)my, 2};
    hid_t sp = H5Screate_simple(4, dims, NULL);
    hid_t ds = H5Dcreate2(fid, name, H5T_IEEE_F64BE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    check(ds, name);
    check(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), name);
    H5Dclose(ds);
    H5Sclose(sp);
}

static void write_int(hid_t fid, const char* name, int v)
{
    hsize_t one = 1;
    hid_t sp = H5Screate_simple(1, &one, NULL);
    hid_t ds = H5Dcreate2(fid, name, H5T_STD_I32LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    check(ds, name);
    check(H5Dwrite(ds, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name);
    H5Dclose(ds);
    H5Sclose(sp);
}

static void write_double(hid_t fid, const char* name, double v)
{
    hsize_t one = 1;
    hid_t sp = H5Screate_simple(1, &one, NULL);
    hid_t ds = H5Dcreate2(fid, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    check(ds, name);
    check(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name);
    H5Dclose(ds);
    H5Sclose(sp);
}

static void write_char(hid_t fid, const char* name, char c)
{
    hsize_t one = 1;
    hid_t sp = H5Screate_simple(1, &one, NULL);
    hid_t ds = H5Dcreate2(fid, name, H5T_FORTRAN_S1, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    check(ds, name);
    check(H5Dwrite(ds, H5T_FORTRAN_S1, H5S_ALL, H5S_ALL, H5P_DEFAULT, &c), name);
    H5Dclose(ds);
    H5Sclose(sp);
}

static void write_array(hid_t fid, const char* name, int rank, const hsize_t* dims, const double* buf)
{
    hid_t sp = H5Screate_simple(rank, dims, NULL);
    hid_t ds = H5Dcreate2(fid, name, H5T_IEEE_F64LE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    check(ds, name);
    check(H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf), name);
    H5Dclose(ds);
    H5Sclose(sp);
}

/* The header of writeheader_hdf5 (readwrite_hdf5.f90:853-918). */
static void write_header(hid_t fid, int temperature, int mgalx, int my, int mz, int mx1, int with_mx1, double time)
{
    write_int(fid, "nparams", temperature ? 13 : 7);
    write_int(fid, "numerop", 1);
    write_int(fid, "mgalx", mgalx);
    write_int(fid, "my", my);
    write_int(fid, "mz", mz);
    if (with_mx1) write_int(fid, "mx1", mx1);   /* older files lack it (readwrite_hdf5.f90:824-826) */
    write_double(fid, "alp", 1.0);
    write_double(fid, "aspect", 1.5);
    write_double(fid, "time", time);
    write_double(fid, "fnu", 0.00125);
    write_double(fid, "cfl", 0.5);
    write_double(fid, "dtfixed", 0.01);
    write_double(fid, "cmassflow", 1.0);
    write_double(fid, "SLIP-VELOCITY", 0.0);
    write_double(fid, "DIVERGENCE", 1.0e-12);
    if (!temperature) return;
    write_double(fid, "alphat", 0.25);
    write_double(fid, "gravx", 0.0);
    write_double(fid, "gravy", -1.0);
    write_double(fid, "gravz", 0.0);
    write_double(fid, "fkappa", 0.0014);
    write_double(fid, "qsource", 2.0);
    write_char(fid, "boundy1", 'd');
    write_char(fid, "boundy2", 'n');
    write_char(fid, "boundz1", 'd');
    write_char(fid, "boundz2", 'n');
    /* bcy(mz,2,0:mx1) and bcz(my,2,0:mx1) in Fortran order */
    double* b = malloc(sizeof(double) * 2 * (mx1 + 1) * (my > mz ? my : mz));
    hsize_t dy[3] = {(hsize_t)(mx1 + 1), 2, (hsize_t)mz};
    for (int k = 0; k <= mx1; ++k)
        for (int w = 0; w < 2; ++w)
            for (int z = 0; z < mz; ++z) b[(k * 2 + w) * mz + z] = 100.0 + k * 8.0 + w * 4.0 + z * 0.125;
    write_array(fid, "bcy", 3, dy, b);
    hsize_t dz[3] = {(hsize_t)(mx1 + 1), 2, (hsize_t)my};
    for (int k = 0; k <= mx1; ++k)
        for (int w = 0; w < 2; ++w)
            for (int y = 0; y < my; ++y) b[(k * 2 + w) * my + y] = -200.0 - k * 8.0 - w * 4.0 - y * 0.0625;
    write_array(fid, "bcz", 3, dz, b);
    free(b);
}

/* One file: fields through one open, the header through a second (as the reference does). */
static void make(const char* path, int latest, int temperature, int my, int mz, int nm, int poly)
{
    const char* names[5] = {"u", "v", "w", "p", "t"};
    const int nf = temperature ? 5 : 4;
    hid_t fcpl = H5Pcreate(H5P_FILE_CREATE), fapl = H5Pcreate(H5P_FILE_ACCESS);
    if (latest) {
        check(H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST), "libver");
        check(H5Pset_link_phase_change(fcpl, 64, 48), "phase change");
        check(H5Pset_userblock(fcpl, 512), "userblock");
    }
    hid_t fid = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, fapl);
    check(fid, path);
    double* buf = malloc(sizeof(double) * 2 * (size_t)my * mz * nm);
    const double pi = 4.0 * atan(1.0);
    for (int f = 0; f < nf; ++f) {
        for (int k = 0; k < nm; ++k)
            for (int z = 0; z < mz; ++z)
                for (int y = 0; y < my; ++y)
                    for (int part = 0; part < 2; ++part) {
                        double v;
                        if (poly) {
                            const double yy = -cos(pi * y / (my - 1)), zz = -cos(pi * z / (mz - 1));
                            v = poly_value(f, k, part, yy, zz);
                        } else {
                            v = exact_value(f, k, z, y, part);
                        }
                        buf[(((size_t)k * mz + z) * my + y) * 2 + part] = v;
                    }
        write_field(fid, names[f], my, mz, nm, buf);
    }
    free(buf);
    check(H5Fclose(fid), "close");
    fid = H5Fopen(path, H5F_ACC_RDWR, fapl);
    check(fid, path);
    const int mx1 = nm - 1;
    write_header(fid, temperature, 3 * (mx1 + 1), my, mz, mx1, !poly, poly ? 3.0 : 12.5);
    check(H5Fclose(fid), "close");
    H5Pclose(fcpl);
    H5Pclose(fapl);
}

int main(int argc, char** argv)
{
    const char* dir = argc > 1 ? argv[1] : ".";
    char p[4096];
    snprintf(p, sizeof p, "%s/uvwp_m17.h5", dir);
    make(p, 0, 0, 17, 17, 3, 0);
    snprintf(p, sizeof p, "%s/uvwp_m25_poly.h5", dir);
    make(p, 0, 0, 25, 25, 2, 1);
    snprintf(p, sizeof p, "%s/uvwpt_m9.h5", dir);
    make(p, 1, 1, 9, 9, 2, 0);
    return 0;
}
