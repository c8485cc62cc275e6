// The feature extractor of the drop-in headers end to end: MtxLoad, CalculateFeatures (GPU),
// CalculateTilesExtraFeatures and FeaturesWrite, appending one feature line to argv[2].
#include "sparse_features.h"

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  MTX<long long, double> m(5);
  if (!m.MtxLoad(argv[1])) return 3;
  m.CalculateFeatures();
  m.CalculateTilesExtraFeatures(argv[1]);
  return m.FeaturesWrite(argv[2]) ? 0 : 4;
}
