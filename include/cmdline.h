/* --key=value command-line parsing with the reference's semantics (include/cmdline.h:21-69):
 * get_arg finds "--key" (prefix match), get_argval returns the text after "--key=", the
 * last occurrence winning. */
#ifndef SPOPS_DROPIN_CMDLINE_H
#define SPOPS_DROPIN_CMDLINE_H

#include <cstring>

inline char *get_arg(int argc, char **argv, const char *key) {
  char *val = nullptr;
  const size_t n = std::strlen(key);
  for (int i = 1; i < argc; i++) {
    char *t = argv[i];
    if (std::strncmp(t, "--", 2) != 0) continue;
    if (std::strncmp(t + 2, key, n) != 0) continue;
    val = argv[i];
  }
  return val;
}

inline char *get_argval(int argc, char **argv, const char *key) {
  char *val = nullptr;
  const size_t n = std::strlen(key);
  for (int i = 1; i < argc; i++) {
    char *t = argv[i];
    if (std::strncmp(t, "--", 2) != 0) continue;
    if (std::strncmp(t + 2, key, n) != 0) continue;
    if (t[2 + n] != '=') continue;
    val = t + 3 + n;
  }
  return val;
}

#endif
