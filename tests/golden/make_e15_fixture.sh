#!/bin/bash
# Builds and runs e15_fixture.f90 with the image's flang and leaves e15_flang.txt here.
set -e
cd "$(dirname "$0")"
FC=$(command -v flang || command -v amdflang || echo /opt/rocm/llvm/bin/flang)
tmp=$(mktemp -d)
"$FC" -O0 -o "$tmp/e15_fixture" e15_fixture.f90
(cd "$tmp" && ./e15_fixture)
cp "$tmp/e15_flang.txt" .
rm -rf "$tmp"
