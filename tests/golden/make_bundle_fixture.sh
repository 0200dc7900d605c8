#!/bin/bash
# Builds and runs bundle_fixture.f90 with the image's flang and leaves bundle_flang.dat here.
set -e
cd "$(dirname "$0")"
FC=$(command -v flang || command -v amdflang || echo /opt/rocm/llvm/bin/flang)
tmp=$(mktemp -d)
"$FC" -O0 -o "$tmp/bundle_fixture" bundle_fixture.f90
(cd "$tmp" && ./bundle_fixture)
cp "$tmp/bundle_flang.dat" .
rm -rf "$tmp"
