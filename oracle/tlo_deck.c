/* Runs tea.in decks through the oracle (tlo_run) and prints, per deck, the
 * records of every timestep: step, iteration count and the field summary.
 * Checker tooling: built with the sanitizers by `make -f oracle/sanitize.mk`
 * to run the oracle's C under AddressSanitizer and UndefinedBehaviorSanitizer;
 * never part of the product.  Exits 1 if any deck could not be read, parsed
 * or run, so a sweep whose decks stop parsing does not look clean. */
#include <stdio.h>
#include <stdlib.h>

#include "tl_oracle.h"

static char *read_file(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  char *t = NULL;
  if (fseek(f, 0, SEEK_END) == 0) {
    const long n = ftell(f);
    if (n >= 0 && fseek(f, 0, SEEK_SET) == 0 && (t = malloc((size_t)n + 1)) != NULL) {
      if (fread(t, 1, (size_t)n, f) != (size_t)n) {
        free(t);
        t = NULL;
      } else {
        t[n] = 0;
      }
    }
  }
  fclose(f);
  return t;
}

int main(int argc, char **argv) {
  enum { MAX_RECORDS = 4096 };
  static tlo_step_record rec[MAX_RECORDS];
  int status = 0;
  for (int a = 1; a < argc; ++a) {
    char *text = read_file(argv[a]);
    char msg[512];
    tlo_params p;
    if (!text) {
      fprintf(stderr, "%s: cannot read\n", argv[a]);
      status = 1;
      continue;
    }
    tlo_default_params(&p);
    if (tlo_parse_deck(text, &p, msg, (int)sizeof msg) != 0) {
      printf("%s: %s\n", argv[a], msg);
      free(text);
      status = 1;
      continue;
    }
    const int n = tlo_run(&p, rec, MAX_RECORDS, 0, msg, (int)sizeof msg);
    if (n < 0) {
      printf("%s: %s\n", argv[a], msg);
      status = 1;
    } else {
      for (int i = 0; i < n; ++i)
        printf("%s step %d iterations %d temperature %.17g\n", argv[a], rec[i].step,
               rec[i].iterations, rec[i].has_summary ? rec[i].summary[4] : 0.0);
    }
    free(text);
  }
  return status;
}
