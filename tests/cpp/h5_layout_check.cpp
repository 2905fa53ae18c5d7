// Reads an HDF5 file with the HDF5 C API alone and prints, per root dataset given on the command line,
// "<name> <rank> <rows> <columns> <maxrows unlimited 0/1> <chunk rows> <chunk columns> <int|double>", the properties
// the reference's reader relies on (src/databases/baseHDF5Database.cpp:190-233): two-dimensional, rows unlimited,
// chunks of one row, native int or double.
#include <hdf5.h>

#include <cstdio>

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return 3;
    for (int a = 2; a < argc; ++a) {
        hid_t d = H5Dopen2(f, argv[a], H5P_DEFAULT);
        if (d < 0) {
            printf("%s missing\n", argv[a]);
            continue;
        }
        hid_t s = H5Dget_space(d);
        hsize_t dims[2] = {0, 0}, maxd[2] = {0, 0}, chunk[2] = {0, 0};
        int rank = H5Sget_simple_extent_ndims(s);
        H5Sget_simple_extent_dims(s, dims, maxd);
        hid_t p = H5Dget_create_plist(d);
        if (H5Pget_layout(p) == H5D_CHUNKED) H5Pget_chunk(p, 2, chunk);
        hid_t t = H5Dget_type(d);
        const char* kind = H5Tequal(t, H5T_NATIVE_INT) > 0 ? "int" : (H5Tequal(t, H5T_NATIVE_DOUBLE) > 0 ? "double" : "other");
        printf("%s %d %llu %llu %d %llu %llu %s\n", argv[a], rank, (unsigned long long)dims[0], (unsigned long long)dims[1],
               maxd[0] == H5S_UNLIMITED ? 1 : 0, (unsigned long long)chunk[0], (unsigned long long)chunk[1], kind);
        H5Tclose(t);
        H5Pclose(p);
        H5Sclose(s);
        H5Dclose(d);
    }
    H5Fclose(f);
    return 0;
}
