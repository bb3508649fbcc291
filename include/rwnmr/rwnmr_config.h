/*
 * rwnmr_config.h -- C ABI of the reader of the reference's configuration
 * files (rwnmr.config, uct.config, cpmg.config, pfgse.config,
 * multitau.config, ga.config: "KEY: value" lines).  Reference paths are
 * under src/backend/ of RandomWalk_NMR (B/).
 */
#ifndef RWNMR_CONFIG_H
#define RWNMR_CONFIG_H

#include "rwnmr.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rwnmr_config_file rwnmr_config_file;

rwnmr_status rwnmr_config_create(rwnmr_config_file **out);
void rwnmr_config_destroy(rwnmr_config_file *cfg);
/* Add the keys of one file (*_config::readConfigFile, e.g.
 * B/ConfigFiles/rwnmr_config.cpp:79-145): on every line each ": " separates
 * a key candidate from the rest of the line, its value, untrimmed.  Read the
 * defaults file first, then the user's (rwnmr_config.cpp:15-23): later
 * values of a key win, except where a list is read with append. */
rwnmr_status rwnmr_config_read(rwnmr_config_file *cfg, const char *path);
/* One more value for key, as if read from a file. */
rwnmr_status rwnmr_config_set(rwnmr_config_file *cfg, const char *key, const char *value);
int32_t rwnmr_config_has(const rwnmr_config_file *cfg, const char *key);
/* Last value of key as text / std::stod / integer / (== "true"). */
rwnmr_status rwnmr_config_string(const rwnmr_config_file *cfg, const char *key, char *buf, uint32_t buflen);
rwnmr_status rwnmr_config_double(const rwnmr_config_file *cfg, const char *key, double *value);
rwnmr_status rwnmr_config_int(const rwnmr_config_file *cfg, const char *key, int64_t *value);
rwnmr_status rwnmr_config_bool(const rwnmr_config_file *cfg, const char *key, int32_t *value);
/* Numeric list "{a, b, ...}" (or one number): the last value of key, or with
 * append every value read, concatenated in reading order (RHO, whose
 * reader pushes onto the list for every file read:
 * rwnmr_config.cpp:178-202).  *n gets the count; out may be NULL to size. */
rwnmr_status rwnmr_config_list(const rwnmr_config_file *cfg, const char *key, int32_t append, double *out,
                               uint32_t capacity, uint32_t *n);

#ifdef __cplusplus
}
#endif

#endif /* RWNMR_CONFIG_H */
