#!/bin/bash
# Compile the reference's native library (common/math/wrapper/wrapper.cpp +
# EVCTCD/CTCD.cpp, built as in the reference README.md:15) from the sources
# where they lie under /root/reference into oracle/_ref/libref.so.  Used only
# by tests that pin the oracle; nothing is copied into this repository.
set -e
REF=${FASTIPC_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=$REF/common/math/wrapper
if [ ! -f "$SRC/wrapper.cpp" ]; then
  echo "reference sources not found under $SRC; skipping" >&2
  exit 0
fi
mkdir -p "$HERE/_ref"
g++ "$SRC/wrapper.cpp" "$SRC/EVCTCD/CTCD.cpp" -o "$HERE/_ref/libref.so" \
    -fPIC -O2 -shared -std=c++1z -mavx2 -mfma -I "$SRC" -w
echo "built $HERE/_ref/libref.so"
