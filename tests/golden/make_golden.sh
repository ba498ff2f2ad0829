#!/usr/bin/env bash
# Regenerates the golden fixtures under tests/golden/ref_<variant>/ by running
# the reference program built from its own sources (oracle/ref_build.sh) in
# scratch directories.  Only output files the reference itself writes are kept:
# cf*.dat result tables (cfph and cfnu where the variant writes them),
# phi_sites.dat, the locnem*.dat files and the stdout timing lines.
# Needs /root/reference (present in the build container, not on GPU hosts).
set -euo pipefail
HERE=$(cd "$(dirname "$0")" && pwd)
REPO=$(cd "$HERE/../.." && pwd)
# usage: make_golden.sh [variant ...]     (default: all variants)
variants=(hot_init_l8 hot_init_l6h4 str_init_l12 pt_l8_stat local_l8 local_frozen pt_frozen pt_l16_c2
          pt_dyn_l6h4 local_dyn_l6h4 local_indep_init local_carry_init)
[ $# -gt 0 ] && variants=("$@")
"$REPO/oracle/ref_build.sh" "${variants[@]}"
for v in "${variants[@]}"; do
    tmp=$(mktemp -d /tmp/j1j2gold.XXXXXX)
    ( cd "$tmp" && "$REPO/oracle/_ref/j1j2_$v" > stdout.txt 2>&1 )
    out="$HERE/ref_$v"
    mkdir -p "$out"
    for f in "$tmp"/cfen*.dat "$tmp"/cfma*.dat "$tmp"/cfst*.dat "$tmp"/cfnm*.dat "$tmp"/cfpt*.dat \
             "$tmp"/cfhs*.dat "$tmp"/cfxy*.dat "$tmp"/cfnd*.dat "$tmp"/cfph*.dat "$tmp"/cfnu*.dat \
             "$tmp"/phi_sites.dat "$tmp"/locnem*.dat; do
        [ -f "$f" ] && cp "$f" "$out/"
    done
    grep '^ *timing' "$tmp/stdout.txt" > "$out/timing.txt" || true
    rm -rf "$tmp"
    echo "golden: $out"
done
