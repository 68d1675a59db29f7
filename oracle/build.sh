#!/bin/bash
# Builds the CPU oracle (test infrastructure) into oracle/_build/.
# FP32 mirror and FP64 oracle from the same source; no FMA contraction, as the
# reference's Julia code emits none.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
out="$here/_build"
mkdir -p "$out"
CFLAGS="-O2 -fPIC -shared -std=c11 -ffp-contract=off -fno-fast-math -fopenmp -Wall -Wno-unused-function"
gcc $CFLAGS -o "$out/liblb_oracle_f32.so.tmp" "$here/lb_oracle.c" -lm
gcc $CFLAGS -DORC_F64 -o "$out/liblb_oracle_f64.so.tmp" "$here/lb_oracle.c" -lm
mv "$out/liblb_oracle_f32.so.tmp" "$out/liblb_oracle_f32.so"
mv "$out/liblb_oracle_f64.so.tmp" "$out/liblb_oracle_f64.so"
