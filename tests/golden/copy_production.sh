#!/usr/bin/env bash
# Copies result tables of the reference's own production runs (data files the
# reference repository holds, not outputs of a build here) for the
# statistical GPU tests, each with its run_meta.txt:
#   tests/golden/ref_prod_clean/  <- data/raw/corr_random_field/p0/c3/
#       J2-1_P0.00_RAN_STRX_s100_scanclean_no_field/L{20,40}: LOCAL, phi = 0,
#       STR-X start, T = 4.00..4.40 step 0.04, INDEP, NCONF = 100,
#       NEQ = NMESS = 100000 sweeps (cfst)
#   tests/golden/ref_prod_pt/     <- data/raw/corr_random_field/p2/c3/
#       J2-1_P0.00_RAN_STRX_s100_scanpt_prod_main/L{20,24,40}: PT SERIAL,
#       phi = 2, xi = 3, 20 replicas over T = 8.0..11.5, STR-X start,
#       NCONF = 1000, NEQ = NMESS = 100000 cycles, PTSWAPEVER = 5
#       (cfen, cfma, cfst, cfnm)
#   tests/golden/ref_prod_pt_hot/ <- the same with HOT starts
#       (J2-1_P0.00_RAN_HOT_s100_scanpt_prod_main/L{20,24}; L = 24: 900 configs)
# Needs /root/reference (present in the build container, not on GPU hosts).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
RAW=${J1J2_REFERENCE:-/root/reference}/data/raw/corr_random_field
copy() {   # source dir, output dir, sizes, tables
    local src=$1 out=$2 sizes=$3 tables=$4
    mkdir -p "$out"
    for L in $sizes; do
        tag=$(printf '%04d%04d' "$L" "$L")
        for t in $tables; do
            cp "$src/L$L/$t$tag.dat" "$out/$t$tag.dat"
        done
        cp "$src/L$L/run_meta.txt" "$out/run_meta_L$L.txt"
    done
    chmod 644 "$out"/*
    echo "copied into $out"
}
copy "$RAW/p0/c3/J2-1_P0.00_RAN_STRX_s100_scanclean_no_field" "$HERE/ref_prod_clean" "20 40" "cfst"
copy "$RAW/p2/c3/J2-1_P0.00_RAN_STRX_s100_scanpt_prod_main" "$HERE/ref_prod_pt" "20 24 40" "cfen cfma cfst cfnm"
copy "$RAW/p2/c3/J2-1_P0.00_RAN_HOT_s100_scanpt_prod_main" "$HERE/ref_prod_pt_hot" "20 24" "cfen cfma cfst cfnm"
