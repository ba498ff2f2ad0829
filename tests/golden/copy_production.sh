#!/usr/bin/env bash
# Copies result tables of the reference's own production runs (data files the
# reference repository holds, not outputs of a build here) into
# tests/golden/ref_prod_clean/ for the statistical GPU tests:
#   data/raw/corr_random_field/p0/c3/J2-1_P0.00_RAN_STRX_s100_scanclean_no_field/L{20,40}
# (LOCAL sampler, phi = 0, STR-X start, T = 4.00..4.40 step 0.04, TEMPPROTOCOL
# INDEP, NCONF = 100, NEQ = NMESS = 100000; run_meta.txt beside each table).
# Needs /root/reference (present in the build container, not on GPU hosts).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
SRC=${J1J2_REFERENCE:-/root/reference}/data/raw/corr_random_field/p0/c3/J2-1_P0.00_RAN_STRX_s100_scanclean_no_field
OUT="$HERE/ref_prod_clean"
mkdir -p "$OUT"
for L in 20 40; do
    tag=$(printf '%04d%04d' "$L" "$L")
    cp "$SRC/L$L/cfst$tag.dat" "$OUT/cfst$tag.dat"
    cp "$SRC/L$L/run_meta.txt" "$OUT/run_meta_L$L.txt"
done
echo "copied into $OUT"
