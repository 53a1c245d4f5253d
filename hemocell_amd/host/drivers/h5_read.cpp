// h5_read -- prints one dataset (or the root attributes) of an HDF5 file written
// by HemoCell::writeOutput as JSON, for checking output files without h5py.
//   h5_read <file.h5> <dataset>   -> {"dims": [...], "data": [...]} (values as float64)
//   h5_read <file.h5> --attrs     -> {"name": [values], ...} of the root attributes
#include <hdf5.h>
#include <hdf5_hl.h>

#include <cstdio>
#include <string>
#include <vector>

static herr_t attr_cb(hid_t loc, const char* name, const H5A_info_t*, void* first) {
  hid_t a = H5Aopen(loc, name, H5P_DEFAULT);
  hid_t sp = H5Aget_space(a);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<double> v((size_t)n);
  H5Aread(a, H5T_NATIVE_DOUBLE, v.data());
  std::printf("%s\"%s\": [", *(int*)first ? "" : ", ", name);
  for (hssize_t i = 0; i < n; ++i) std::printf("%s%.17g", i ? ", " : "", v[(size_t)i]);
  std::printf("]");
  *(int*)first = 0;
  H5Sclose(sp);
  H5Aclose(a);
  return 0;
}

int main(int argc, char* argv[]) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <file.h5> <dataset|--attrs>\n", argv[0]);
    return 2;
  }
  hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
  if (f < 0) return 1;
  if (std::string(argv[2]) == "--attrs") {
    int first = 1;
    std::printf("{");
    H5Aiterate2(f, H5_INDEX_NAME, H5_ITER_INC, nullptr, attr_cb, &first);
    std::printf("}\n");
    H5Fclose(f);
    return 0;
  }
  hid_t d = H5Dopen2(f, argv[2], H5P_DEFAULT);
  if (d < 0) return 1;
  hid_t sp = H5Dget_space(d);
  const int rank = H5Sget_simple_extent_ndims(sp);
  std::vector<hsize_t> dims((size_t)rank);
  H5Sget_simple_extent_dims(sp, dims.data(), nullptr);
  const hssize_t n = H5Sget_simple_extent_npoints(sp);
  std::vector<double> v((size_t)n);
  if (n > 0) H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data());
  std::printf("{\"dims\": [");
  for (int i = 0; i < rank; ++i) std::printf("%s%llu", i ? ", " : "", (unsigned long long)dims[(size_t)i]);
  std::printf("], \"data\": [");
  for (hssize_t i = 0; i < n; ++i) std::printf("%s%.9g", i ? ", " : "", v[(size_t)i]);
  std::printf("]}\n");
  H5Sclose(sp);
  H5Dclose(d);
  H5Fclose(f);
  return 0;
}
