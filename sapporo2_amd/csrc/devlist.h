// devlist.h -- the reference's device-list config file (./sapporo2.config:
// first line ndev, then one device id per line; testCodes/sapporo2.config.README).
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace sap {

// Returns the device to use: the first listed id when ndev > 0, else 0
// (ndev <= 0 means "let the library choose"; one GPU is driven).
inline int device_from_list(int ndev, const int* list) {
  if (ndev > 0 && list) return list[0];
  return 0;
}

inline int device_from_config_file(const char* path = "sapporo2.config") {
  FILE* fd = fopen(path, "r");
  if (!fd) {
    fprintf(stderr, " sapporo2_amd::open - no config file is found \n");
    return 0;
  }
  fprintf(stderr, "sapporo2_amd::open - config file is found\n");
  char line[256];
  int how_many = 0, dev = 0;
  if (fgets(line, sizeof line, fd)) how_many = atoi(line);
  if (how_many > 0 && fgets(line, sizeof line, fd)) dev = atoi(line);
  fclose(fd);
  return dev;
}

// Working precision from SAPPORO2_AMD_PRECISION ("mixed"/"ds" or "double"/"fp64").
inline int precision_from_env(int dflt) {
  const char* p = getenv("SAPPORO2_AMD_PRECISION");
  if (!p) return dflt;
  if (!strcmp(p, "double") || !strcmp(p, "fp64") || !strcmp(p, "2")) return 2;
  if (!strcmp(p, "mixed") || !strcmp(p, "ds") || !strcmp(p, "1")) return 1;
  return dflt;
}

}  // namespace sap
