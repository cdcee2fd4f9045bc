/* Reader for the scene-spec container written by slimtracincpp_amd/specfile.py.
 *
 * Test infrastructure: used by the CPU oracle (oracle/slim_oracle.c) and by the driver that runs the
 * compiled reference (oracle/ref_driver.cpp).  Header-only C, also valid C++.
 */
#ifndef SLIM_ORACLE_SPECFILE_H
#define SLIM_ORACLE_SPECFILE_H

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    uint8_t *data;
    size_t size;
    uint32_t count;
} SpecFile;

typedef struct {
    const void *ptr;
    uint32_t dtype;
    uint32_t count;
} SpecSection;

static int spec_open(SpecFile *sf, const char *path) {
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    fseek(f, 0, SEEK_END);
    long n = ftell(f);
    fseek(f, 0, SEEK_SET);
    sf->data = (uint8_t *)malloc((size_t)n);
    sf->size = (size_t)n;
    if (!sf->data || fread(sf->data, 1, (size_t)n, f) != (size_t)n) { fclose(f); return -2; }
    fclose(f);
    if (n < 16 || memcmp(sf->data, "SLMSPEC1", 8) != 0) return -3;
    memcpy(&sf->count, sf->data + 8, 4);
    return 0;
}

static void spec_close(SpecFile *sf) {
    free(sf->data);
    sf->data = NULL;
}

static size_t spec_elem_size(uint32_t dtype) { return dtype == 3 ? 1 : 4; }

/* Returns 1 and fills *out when the section exists, else 0. */
static int spec_find(const SpecFile *sf, const char *name, SpecSection *out) {
    size_t off = 16;
    for (uint32_t i = 0; i < sf->count; i++) {
        if (off + 32 > sf->size) return 0;
        const char *nm = (const char *)(sf->data + off);
        uint32_t dtype, count;
        memcpy(&dtype, sf->data + off + 24, 4);
        memcpy(&count, sf->data + off + 28, 4);
        size_t bytes = (size_t)count * spec_elem_size(dtype);
        size_t padded = (bytes + 7) & ~(size_t)7;
        if (strncmp(nm, name, 24) == 0) {
            out->ptr = sf->data + off + 32;
            out->dtype = dtype;
            out->count = count;
            return 1;
        }
        off += 32 + padded;
    }
    return 0;
}

static const void *spec_get(const SpecFile *sf, const char *name, uint32_t *count) {
    SpecSection s;
    if (!spec_find(sf, name, &s)) { if (count) *count = 0; return NULL; }
    if (count) *count = s.count;
    return s.ptr;
}

#endif
