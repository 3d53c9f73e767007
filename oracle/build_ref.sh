#!/bin/bash
# Compile the reference's native sampler (bor_stream_kernel.c, plain C + libm)
# from where it lies into oracle/_ref/ (git-ignored).  Used only as a checker.
set -e
REF=${GHOST_REFERENCE:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
mkdir -p "$HERE/_ref"
gcc -O2 -shared -fPIC -o "$HERE/_ref/bor_stream_kernel_ref.so" "$REF/bor_stream_kernel.c" -lm
echo "$HERE/_ref/bor_stream_kernel_ref.so"
