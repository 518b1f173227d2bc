#!/bin/sh
# Builds the CPU oracle (test infrastructure only) into oracle/:
#   liboracle_f32.so  -- `real` = float, the reference's FP32 arithmetic
#   liboracle_f64.so  -- `real` = double, the FP64 yardstick
# -ffp-contract=off keeps every expression rounded as written; x86-64-v3 keeps
# the library loadable on any AVX2 host (the bench host may differ from this one).
set -e
cd "$(dirname "$0")"
CC=${CC:-gcc}
CFLAGS="-O3 -march=x86-64-v3 -fopenmp -fPIC -shared -std=c11 -ffp-contract=off -Wall -Wno-unused-function"
$CC $CFLAGS sph_oracle.c -o liboracle_f32.so.tmp -lm
mv -f liboracle_f32.so.tmp liboracle_f32.so
$CC $CFLAGS -DORACLE_F64 sph_oracle.c -o liboracle_f64.so.tmp -lm
mv -f liboracle_f64.so.tmp liboracle_f64.so
$CC $CFLAGS surface_oracle.c -o liboracle_surface.so.tmp -lm
mv -f liboracle_surface.so.tmp liboracle_surface.so
