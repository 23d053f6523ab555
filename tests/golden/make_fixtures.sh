#!/bin/sh
# Copies the data files the reference's own tests read for this path
# (tests/test_core_freeze.py: updc/cache_abi.json, updc/core_freeze.json) into
# this directory.  They are data, not source: the LinkMap ABI constants and the
# SHA-256 of the frozen v17 C core that the oracle restates.
set -e
REF=${1:-/root/reference}
HERE=$(cd "$(dirname "$0")" && pwd)
cp "$REF/updc/cache_abi.json" "$HERE/cache_abi.json"
cp "$REF/updc/core_freeze.json" "$HERE/core_freeze.json"
