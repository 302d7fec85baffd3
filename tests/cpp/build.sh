#!/bin/bash
# Builds tests/cpp/tacs_cube_driver (a C++ driver written against the
# reference's class names, compiled against include/TACSHip.h and linked with
# tacs_amd/lib/libtacs_hip.so and MPICH) for tests/test_gpu_cpp_driver.py.
set -e
HERE=$(cd "$(dirname "$0")" && pwd)
ROOT=$(cd "$HERE/../.." && pwd)
MPI=${TACS_MPI_PREFIX:-/opt/conda}
/opt/rocm/bin/hipcc -O2 -std=c++17 -I "$ROOT/include" -idirafter "$MPI/include" \
  "$HERE/tacs_cube_driver.cpp" -o "$HERE/tacs_cube_driver" \
  -L "$ROOT/tacs_amd/lib" -ltacs_hip -Wl,"$MPI/lib/libmpi.so.12" \
  -Wl,-rpath,'$ORIGIN/../../tacs_amd/lib' -Wl,-rpath,/usr/lib/x86_64-linux-gnu \
  -Wl,-rpath,"$MPI/lib"
# (the system directory comes before MPICH's in the search path: MPICH's
# prefix also holds an older libstdc++ than the one the library needs)
echo "built $HERE/tacs_cube_driver"
