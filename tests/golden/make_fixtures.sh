#!/bin/bash
# Regenerates tests/golden/data/*.npy from the reference's own test fixtures
# (test/data/O3.jld, test/data/linalg.jld). Runs only where the reference and
# libhdf5 are present; the outputs are committed so nothing else needs them.
set -euo pipefail
here="$(cd "$(dirname "$0")" && pwd)"
ref="${REFERENCE_ROOT:-/root/reference}"
h5="${HDF5_PREFIX:-/opt/conda}"
tmp="$(mktemp -d)"
trap 'rm -rf "$tmp"' EXIT
gcc -O1 -std=c99 "$here/extract_fixtures.c" -I"$h5/include" -L"$h5/lib" -lhdf5 \
    -Wl,-rpath,"$h5/lib" -o "$tmp/extract_fixtures"
out="$here/data"
mkdir -p "$out"
"$tmp/extract_fixtures" "$ref/test/data/O3.jld" "$out" \
  randconf afterlocal_hsfield greens Ul Dl Tl Ur Dr Tr delta_i M \
  afterupdate_greens afterlocal_greens Bplus Bminus Bplus_CBFalse Bminus_CBFalse \
  nob_eVplus nob_eVminus eVplus eVminus eVexpop nob_eVexpop \
  nob_chkr_hop1 nob_chkr_hop2 nob_chkr_hop_inv1 nob_chkr_hop_inv2 \
  nob_chkr_hop_half1 nob_chkr_hop_half2 nob_chkr_hop_half_inv1 nob_chkr_hop_half_inv2 \
  nob_chkr_mu nob_chkr_mu_half nob_chkr_mu_inv nob_chkr_mu_half_inv \
  chkr_hop1 chkr_hop2 chkr_hop_inv1 chkr_hop_inv2 chkr_hop_half1 chkr_hop_half2 \
  chkr_hop_half_inv1 chkr_hop_half_inv2 chkr_mu chkr_mu_inv \
  hopping_matrix_exp hopping_matrix_exp_inv peirls11 peirls12 peirls21 peirls22 \
  > /dev/null
"$tmp/extract_fixtures" "$ref/test/data/linalg.jld" "$out" B_QR_U B_QR_D B_QR_T > /dev/null
python3 "$here/write_manifest.py" "$out"
