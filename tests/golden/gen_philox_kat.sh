#!/bin/sh
# Regenerates tests/golden/philox_kat.json from ROCm's rocrand Philox (host build).
set -e
here=$(dirname "$0")
tmp=$(mktemp -d)
/opt/rocm/bin/hipcc -O1 "$here/gen_philox_kat.cpp" -o "$tmp/kat"
"$tmp/kat" > "$here/philox_kat.json"
rm -rf "$tmp"
