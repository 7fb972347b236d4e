// Driver over the reference's own type and fusion-rule sources. TEST
// INFRASTRUCTURE ONLY: built by oracle/Makefile from the reference sources in
// place (util/promote.hpp, util/DataType.cpp, util/VariantType.cpp,
// util/ReductionType.cpp, util/BinaryType.cpp, runtime/Pattern.cpp) into
// oracle/_ref/ref_tables, whose JSON output pins the tables the backend and
// the oracle restate.
#include "runtime/Pattern.hpp"
#include "util/BinaryType.hpp"
#include "util/ReductionType.hpp"
#include "util/promote.hpp"
#include <cstdio>

using namespace map::detail;

int main() {
  std::printf("{\"promote\": [");
  for (int a = 1; a < N_DATATYPE; a++) {
    std::printf("%s[", a > 1 ? ", " : "");
    for (int b = 1; b < N_DATATYPE; b++)
      std::printf("%s%d", b > 1 ? ", " : "", (int)promote(DataType((DataTypeEnum)a), DataType((DataTypeEnum)b)).get());
    std::printf("]");
  }
  std::printf("],\n\"neutral\": {");
  bool first = true;
  for (int r : {SUM, PROD, rAND, rOR, MAX, MIN})
    for (int d = 1; d < N_DATATYPE; d++) {
      VariantType v = ReductionType((ReductionEnum)r).neutral(DataType((DataTypeEnum)d));
      std::printf("%s\"%d,%d\": \"%s\"", first ? "" : ", ", r, d, v.toString().c_str());
      first = false;
    }
  std::printf("},\n\"binary_code\": {");
  first = true;
  for (int b = 1; b < N_BINARY; b++) {
    if (b == MARK_BINARY) continue;
    std::printf("%s\"%d\": \"%s\"", first ? "" : ", ", b, BinaryType((BinaryEnum)b).code().c_str());
    first = false;
  }
  std::printf("},\n\"float_literal\": {");
  const float samples[] = {0.1f, -0.25f, 1.0f / 3.0f, 0.015625f, 123.456789f, -7.0000005f, 1e-7f, 2.5e-6f};
  first = true;
  for (float f : samples) {
    std::printf("%s\"%.9g\": \"%s\"", first ? "" : ", ", (double)f, VariantType(f).toString().c_str());
    first = false;
  }
  std::printf("},\n\"fusion\": [");
  const int pats[] = {FREE, LOCAL, FOCAL, ZONAL, RADIAL, SPREAD, SPECIAL, BARRIER, FREE | LOCAL, FOCAL | LOCAL,
                      FOCAL | LOCAL | FREE, ZONAL | LOCAL, ZONAL | LOCAL | FREE, FOCAL | ZONAL | LOCAL};
  first = true;
  for (int a : pats)
    for (int b : pats) {
      Pattern pa, pb;
      pa.pat = (PatternEnum)a;
      pb.pat = (PatternEnum)b;
      std::printf("%s[%d, %d, %d, %d]", first ? "" : ", ", a, b, (int)canPipeFuse(pa, pb), (int)canFlatFuse(pa, pb));
      first = false;
    }
  std::printf("]}\n");
  return 0;
}
