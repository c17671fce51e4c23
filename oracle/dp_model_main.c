/* Command-line driver of the protocol model for the ThreadSanitizer build
 * (test infrastructure): dp_model_tsan groups resident lock bar skip barriers weak
 * prints "P=<p> violations=<v> deadlock=<d>". */
#include <stdio.h>
#include <stdlib.h>
#include "oracle.h"

int main(int argc, char** argv) {
  if (argc != 8) {
    fprintf(stderr, "usage: %s groups resident lock bar skip barriers weak\n", argv[0]);
    return 2;
  }
  const int groups = atoi(argv[1]);
  int* ids = (int*)calloc((size_t)(groups > 0 ? groups : 1), sizeof(int));
  int p = 0;
  long long bad = 0;
  orc_dp_model_set_weak(atoi(argv[7]));
  const int dead = orc_dp_model(groups, atoi(argv[2]), atoi(argv[3]), atoi(argv[4]), atoi(argv[5]), atoi(argv[6]),
                                5000, &p, ids, &bad);
  printf("P=%d violations=%lld deadlock=%d\n", p, bad, dead);
  free(ids);
  return 0;
}
