// tune.h -- process-wide kernel-variant switches (see tune.cpp).
#pragma once

namespace phx {
bool tune_set(const char *key, int value);
bool tune_get(const char *key, int *value);
int tune_value(const char *key);
}  // namespace phx
