#!/bin/bash
# Builds the reference's own strategy plugin (strat.so) from the table
# headers where they lie under /root/reference/strat, into oracle/_ref/.
# It mirrors strat/proc_dir-lnx.sh + strat/proc_all-lnx.sh (gcc -fPIC on a
# wrapper that includes every NNNNN.h of a family, then ld -shared), but
# writes nothing under /root/reference.  Used only by the tests, to check
# this library's generated tables against the reference's.
set -e
REF=${1:-/root/reference}
OUT=$(cd "$(dirname "$0")" && pwd)/_ref
[ -d "$REF/strat" ] || { echo "no reference at $REF"; exit 0; }
mkdir -p "$OUT"
cd "$OUT"
for fam in mmstep BrentL colcyc cycloc rowcyc cycwor; do
  {
    echo "#include \"$REF/strat/hdr_unx.h\""
    for h in "$REF/strat/$fam"/*.h; do
      echo "EXPORT_VAR"
      echo "#include \"$h\""
    done
  } > "ref_$fam.c"
  gcc -O0 -fPIC -c "ref_$fam.c" -o "ref_$fam.o" &
done
wait
ld -shared -o strat.so ref_*.o
rm -f ref_*.o ref_*.c
echo "built $OUT/strat.so"
