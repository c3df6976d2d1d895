#!/bin/sh
# Rebuilds the HDF5 fixtures of the field-file reader with the HDF5 C library
# (HDF5_DIR, default /opt/conda: headers in include/, libhdf5 in lib/).
set -e
here=$(cd "$(dirname "$0")" && pwd)
H5=${HDF5_DIR:-/opt/conda}
tmp=$(mktemp -d)
gcc -O2 -std=c99 -Wall -I"$H5/include" "$here/make_uvwp_h5.c" -o "$tmp/make_uvwp_h5" \
    -L"$H5/lib" -lhdf5 -lm -Wl,-rpath,"$H5/lib"
"$tmp/make_uvwp_h5" "$here"
rm -r "$tmp"
