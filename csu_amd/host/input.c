/*
 * input.c -- input.dat reader.
 *
 * Grammar of the reference reader (input.c:7-115), kept exactly because case
 * files must mean the same thing to both programs:
 *   - spaces and '\r' are dropped everywhere;
 *   - the first ',' of a line ends the name, the rest is the value;
 *   - a line is committed at '\n' when it had a ','; a line with a name but
 *     no ',' is committed with an empty value;
 *   - '#' ends the current entry without committing it; characters after it
 *     are ignored until a ',' re-enters the value state (the reference's
 *     ',' branch runs in every state), so '#' without a later ',' ends
 *     parsing;
 *   - a last line without '\n' is not committed;
 *   - lookups return the last duplicate (input.c:142-163).
 * Entries are capped at 100 and names/values at 49 characters, the
 * reference's buffer sizes (input.c:37-51); longer text is truncated here
 * instead of overflowing.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "csu_host.h"

#define IN_MAX_ENTRIES 100
#define IN_MAX_CHARS   50

struct csu_input {
    int n;
    char name[IN_MAX_ENTRIES][IN_MAX_CHARS];
    char value[IN_MAX_ENTRIES][IN_MAX_CHARS];
};

int csu_input_read(const char* path, csu_input** out)
{
    FILE* f = fopen(path, "r");
    if (!f) return CSU_ERR_IO;
    csu_input* in = calloc(1, sizeof(*in));
    if (!in) { fclose(f); return CSU_ERR_NOMEM; }
    int in_name = 1, in_value = 0, ii = 0, jj = 0, c;
    while ((c = fgetc(f)) != EOF && ii < IN_MAX_ENTRIES) {
        if (c == ' ' || c == '\r') continue;
        if (c == ',') {
            /* in every state (also after '#' or a first ','), as input.c:63-69:
               terminate the name at the current index and (re)start the value */
            in->name[ii][jj < IN_MAX_CHARS ? jj : IN_MAX_CHARS - 1] = '\0';
            in_name = 0; in_value = 1; jj = 0;
        } else if (c == '#') {
            in->value[ii][jj < IN_MAX_CHARS ? jj : IN_MAX_CHARS - 1] = '\0';
            in_name = 0; in_value = 0;
        } else if (c == '\n') {
            if (in_value) {
                in->value[ii][jj < IN_MAX_CHARS ? jj : IN_MAX_CHARS - 1] = '\0';
                in_name = 1; in_value = 0; ii++; jj = 0;
            }
            if (in_name && jj > 0) {
                in->name[ii][jj < IN_MAX_CHARS ? jj : IN_MAX_CHARS - 1] = '\0';
                ii++; jj = 0;
            }
        } else if (in_name) {
            if (jj < IN_MAX_CHARS - 1) in->name[ii][jj] = (char)c;
            jj++;
        } else if (in_value) {
            if (jj < IN_MAX_CHARS - 1) in->value[ii][jj] = (char)c;
            jj++;
        }
    }
    fclose(f);
    in->n = ii;
    *out = in;
    return CSU_OK;
}

void csu_input_free(csu_input* in) { free(in); }

int csu_input_count(const csu_input* in) { return in ? in->n : 0; }

const char* csu_input_name(const csu_input* in, int i)
{
    return (in && i >= 0 && i < in->n) ? in->name[i] : NULL;
}

const char* csu_input_value(const csu_input* in, int i)
{
    return (in && i >= 0 && i < in->n) ? in->value[i] : NULL;
}

const char* csu_input_get(const csu_input* in, const char* name)
{
    const char* s = NULL;
    if (!in) return NULL;
    for (int i = 0; i < in->n; i++)
        if (strcmp(name, in->name[i]) == 0) s = in->value[i];
    return s;
}
