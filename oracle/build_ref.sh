#!/usr/bin/env bash
# Builds the reference StreamingBC sources, where they lie under
# /root/reference/src, together with oracle/ref_driver.c into
# oracle/_ref/libsbcref.so.  Test infrastructure: used to write the golden
# fixtures and to time the reference's OpenMP update as the CPU baseline.
#
# The reference Makefile compiles the .c files with g++, which fails under
# g++ 11 (pointer/zero comparisons in src/streamingbc_internal.c:1199 and
# src/stinger/stinger-utils.c:72,84), so the same sources and flags
# (-O3 -fopenmp, warnings off) are compiled as C instead.  src/sbcMain.c is
# left out: it holds only the CLI main().
set -euo pipefail
REF=${SBC_REFERENCE_ROOT:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
OUT="$HERE/_ref"
if [ ! -d "$REF/src" ]; then
    echo "build_ref.sh: $REF/src not found; reference build skipped" >&2
    exit 0
fi
mkdir -p "$OUT"
S="$REF/src"
gcc -std=gnu99 -O3 -fopenmp -fPIC -shared -w \
    -I"$S" -I"$S/stinger" \
    "$S/streamingbc.c" "$S/streamingbc_init.c" "$S/streamingbc_internal.c" \
    "$S/streamingbc_internal_fine.c" "$S/streamingbc_aux.c" "$S/sbcMain_utils.c" \
    "$S/stinger/stinger.c" "$S/stinger/stinger-utils.c" "$S/stinger/timer.c" \
    "$S/stinger/xmalloc.c" "$S/stinger/x86-full-empty.c" \
    "$HERE/ref_driver.c" \
    -o "$OUT/libsbcref.so.tmp" -lm -lrt
mv "$OUT/libsbcref.so.tmp" "$OUT/libsbcref.so"
echo "built $OUT/libsbcref.so"
