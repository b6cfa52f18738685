/*
 * ns_config.c -- run configuration store (include/ns_config.h).
 *
 * Behaviour follows the reference's src/common/config_parser.c:
 *   - INI: "[section]" prefixes keys as "section.key"; '#' and ';' start
 *     comment lines; surrounding whitespace is trimmed  (config_parser.c:205-263)
 *   - JSON: nested objects flatten to dotted keys; numbers are stored as
 *     "%.15g" text; true/false/null become "true"/"false"/""  (:286-486)
 *   - later files override earlier ones                      (:543-554)
 *   - schema validation by type, range and allowed strings   (:556-634)
 *
 * Keys understood by ns_spectral_configure (spectral_config.c):
 *   solver.type, grid.nx, grid.ny, domain.Lx, domain.Ly, fluid.Re, time.dt,
 *   time.t_end, forcing.type, forcing.energy_target, forcing.correlation_time,
 *   spectral.hyperviscosity.coefficient, spectral.hyperviscosity.order
 *   (the reference's keys, config/parameters.ini:1-26), plus forcing.kmin,
 *   forcing.kmax, forcing.seed, forcing.injection_rate,
 *   spectral.hyperviscosity.{enabled,cutoff_ratio,sfd_strength,sfd_timescale},
 *   time.cfl_factor, time.adaptive, fluid.drag, spectral.scheme,
 *   spectral.precision.
 */
#include "ns_config.h"

#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { KEY_MAX = 256, LINE_MAX_LEN = 1024 };

static void fail(char *error, size_t n, const char *what, const char *detail) {
    if (!error || n == 0) return;
    if (detail) snprintf(error, n, "%s: %s", what, detail);
    else snprintf(error, n, "%s", what);
}

static char *dup_str(const char *s) {
    const size_t n = strlen(s);
    char *r = (char *)malloc(n + 1);
    if (r) memcpy(r, s, n + 1);
    return r;
}

static int ieq(const char *a, const char *b) {
    for (; *a && *b; ++a, ++b)
        if (tolower((unsigned char)*a) != tolower((unsigned char)*b)) return 0;
    return *a == *b;
}

/* in-place trim; returns the start of the trimmed text inside s */
static char *trim(char *s) {
    while (*s && isspace((unsigned char)*s)) ++s;
    char *e = s + strlen(s);
    while (e > s && isspace((unsigned char)e[-1])) --e;
    *e = '\0';
    return s;
}

static NSConfigEntry *find(const NSConfig *c, const char *key) {
    if (!c || !key) return NULL;
    for (size_t i = 0; i < c->count; ++i)
        if (strcmp(c->entries[i].key, key) == 0) return &c->entries[i];
    return NULL;
}

void ns_config_init(NSConfig *c) {
    if (!c) return;
    c->entries = NULL;
    c->count = 0;
    c->capacity = 0;
}

void ns_config_free(NSConfig *c) {
    if (!c) return;
    for (size_t i = 0; i < c->count; ++i) {
        free(c->entries[i].key);
        free(c->entries[i].value);
    }
    free(c->entries);
    ns_config_init(c);
}

int ns_config_set(NSConfig *c, const char *key, const char *value) {
    if (!c || !key || !value) return -1;
    NSConfigEntry *e = find(c, key);
    char *v = dup_str(value);
    if (!v) return -1;
    if (e) {
        free(e->value);
        e->value = v;
        return 0;
    }
    if (c->count == c->capacity) {
        const size_t cap = c->capacity ? 2 * c->capacity : 16;
        NSConfigEntry *ne = (NSConfigEntry *)realloc(c->entries, cap * sizeof(NSConfigEntry));
        if (!ne) {
            free(v);
            return -1;
        }
        c->entries = ne;
        c->capacity = cap;
    }
    char *k = dup_str(key);
    if (!k) {
        free(v);
        return -1;
    }
    c->entries[c->count].key = k;
    c->entries[c->count].value = v;
    c->count++;
    return 0;
}

int ns_config_has_key(const NSConfig *c, const char *key) { return find(c, key) != NULL; }

const char *ns_config_get_string(const NSConfig *c, const char *key, const char *def) {
    const NSConfigEntry *e = find(c, key);
    return e ? e->value : def;
}

static int parse_long(const char *s, long *out) {
    if (!s || !*s) return 0;
    char *end = NULL;
    long v = strtol(s, &end, 10);
    if (end && *end) return 0;
    *out = v;
    return 1;
}

static int parse_double(const char *s, double *out) {
    if (!s || !*s) return 0;
    char *end = NULL;
    double v = strtod(s, &end);
    if (end && *end) return 0;
    *out = v;
    return 1;
}

/* 1 true, 0 false, -1 not a boolean */
static int parse_bool(const char *s) {
    static const char *yes[] = {"true", "yes", "on"}, *no[] = {"false", "no", "off"};
    if (!s || !*s) return -1;
    if (strcmp(s, "1") == 0) return 1;
    if (strcmp(s, "0") == 0) return 0;
    for (int i = 0; i < 3; ++i) {
        if (ieq(s, yes[i])) return 1;
        if (ieq(s, no[i])) return 0;
    }
    return -1;
}

int ns_config_get_int(const NSConfig *c, const char *key, int def) {
    long v;
    const NSConfigEntry *e = find(c, key);
    return (e && parse_long(e->value, &v)) ? (int)v : def;
}

double ns_config_get_double(const NSConfig *c, const char *key, double def) {
    double v;
    const NSConfigEntry *e = find(c, key);
    return (e && parse_double(e->value, &v)) ? v : def;
}

int ns_config_get_bool(const NSConfig *c, const char *key, int def) {
    const NSConfigEntry *e = find(c, key);
    if (!e) return def;
    const int b = parse_bool(e->value);
    return b < 0 ? def : b;
}

/* ------------------------------------------------------------------ INI */

static int load_ini(NSConfig *c, FILE *fp, char *error, size_t n) {
    char section[KEY_MAX] = "";
    char buf[LINE_MAX_LEN];
    while (fgets(buf, sizeof(buf), fp)) {
        char *line = trim(buf);
        if (!*line || *line == '#' || *line == ';') continue;
        if (*line == '[') {
            char *close = strchr(line, ']');
            if (!close) {
                fail(error, n, "Malformed section header in INI file", line);
                return -1;
            }
            *close = '\0';
            snprintf(section, sizeof(section), "%s", line + 1);
            char *t = trim(section);
            memmove(section, t, strlen(t) + 1);
            continue;
        }
        char *eq = strchr(line, '=');
        if (!eq) {
            fail(error, n, "Malformed key-value line in INI file", line);
            return -1;
        }
        *eq = '\0';
        char kbuf[KEY_MAX], vbuf[512];
        snprintf(kbuf, sizeof(kbuf), "%s", line);
        snprintf(vbuf, sizeof(vbuf), "%s", eq + 1);
        const char *k = trim(kbuf), *v = trim(vbuf);
        char full[KEY_MAX];
        int w = section[0] ? snprintf(full, sizeof(full), "%s.%s", section, k)
                           : snprintf(full, sizeof(full), "%s", k);
        if (w < 0 || (size_t)w >= sizeof(full)) {
            fail(error, n, "Combined key too long", k);
            return -1;
        }
        if (ns_config_set(c, full, v) != 0) {
            fail(error, n, "Failed to store configuration key", full);
            return -1;
        }
    }
    return 0;
}

/* ----------------------------------------------------------------- JSON */

typedef struct {
    const char *p;
    char *error;
    size_t n;
} JsonCursor;

static void ws(JsonCursor *j) {
    while (*j->p && isspace((unsigned char)*j->p)) j->p++;
}

/* parse a JSON string at the cursor into a fresh buffer */
static char *jstring(JsonCursor *j) {
    ws(j);
    if (*j->p != '"') {
        fail(j->error, j->n, "Expected string in JSON configuration", NULL);
        return NULL;
    }
    const char *s = ++j->p;
    size_t cap = 32, len = 0;
    char *out = (char *)malloc(cap);
    if (!out) return NULL;
    while (*s && *s != '"') {
        char ch = *s++;
        if (ch == '\\') {
            if (!*s) {
                free(out);
                fail(j->error, j->n, "Incomplete escape sequence in JSON string", NULL);
                return NULL;
            }
            const char e = *s++;
            switch (e) {
            case 'b': ch = '\b'; break;
            case 'f': ch = '\f'; break;
            case 'n': ch = '\n'; break;
            case 'r': ch = '\r'; break;
            case 't': ch = '\t'; break;
            default: ch = e; break; /* \" \\ \/ and anything else: the character */
            }
        }
        if (len + 2 > cap) {
            cap *= 2;
            char *t = (char *)realloc(out, cap);
            if (!t) {
                free(out);
                return NULL;
            }
            out = t;
        }
        out[len++] = ch;
    }
    if (*s != '"') {
        free(out);
        fail(j->error, j->n, "Unterminated string in JSON configuration", NULL);
        return NULL;
    }
    out[len] = '\0';
    j->p = s + 1;
    return out;
}

static int jvalue(NSConfig *c, JsonCursor *j, const char *key);

static int jobject(NSConfig *c, JsonCursor *j, const char *prefix) {
    ws(j);
    if (*j->p != '{') {
        fail(j->error, j->n, "Expected '{' in JSON configuration", NULL);
        return -1;
    }
    j->p++;
    ws(j);
    if (*j->p == '}') {
        j->p++;
        return 0;
    }
    for (;;) {
        char *name = jstring(j);
        if (!name) return -1;
        ws(j);
        if (*j->p != ':') {
            free(name);
            fail(j->error, j->n, "Expected ':' in JSON key/value pair", NULL);
            return -1;
        }
        j->p++;
        char full[KEY_MAX];
        int w = (prefix && *prefix) ? snprintf(full, sizeof(full), "%s.%s", prefix, name)
                                    : snprintf(full, sizeof(full), "%s", name);
        free(name);
        if (w < 0 || (size_t)w >= sizeof(full)) {
            fail(j->error, j->n, "Combined JSON key too long", NULL);
            return -1;
        }
        if (jvalue(c, j, full) != 0) return -1;
        ws(j);
        if (*j->p == ',') {
            j->p++;
            continue;
        }
        if (*j->p == '}') {
            j->p++;
            return 0;
        }
        fail(j->error, j->n, *j->p ? "Expected ',' or '}' in JSON object"
                                   : "Unexpected end of JSON content", NULL);
        return -1;
    }
}

static int jstore(NSConfig *c, JsonCursor *j, const char *key, const char *value) {
    if (ns_config_set(c, key, value) != 0) {
        fail(j->error, j->n, "Failed to store JSON value", key);
        return -1;
    }
    return 0;
}

static int jvalue(NSConfig *c, JsonCursor *j, const char *key) {
    ws(j);
    const char ch = *j->p;
    if (!ch) {
        fail(j->error, j->n, "Unexpected end of JSON value", NULL);
        return -1;
    }
    if (ch == '{') return jobject(c, j, key);
    if (ch == '"') {
        char *s = jstring(j);
        if (!s) return -1;
        const int rc = jstore(c, j, key, s);
        free(s);
        return rc;
    }
    static const struct { const char *word, *stored; } lit[] = {
        {"true", "true"}, {"false", "false"}, {"null", ""}};
    for (int i = 0; i < 3; ++i) {
        const size_t L = strlen(lit[i].word);
        if (strncmp(j->p, lit[i].word, L) == 0) {
            j->p += L;
            return jstore(c, j, key, lit[i].stored);
        }
    }
    char *end = NULL;
    const double num = strtod(j->p, &end);
    if (end == j->p) {
        fail(j->error, j->n, "Invalid JSON value", key);
        return -1;
    }
    j->p = end;
    char text[64];
    snprintf(text, sizeof(text), "%.15g", num);
    return jstore(c, j, key, text);
}

static int load_json(NSConfig *c, const char *text, char *error, size_t n) {
    JsonCursor j = {text, error, n};
    if (jobject(c, &j, "") != 0) return -1;
    ws(&j);
    if (*j.p) {
        fail(error, n, "Unexpected trailing data in JSON file", NULL);
        return -1;
    }
    return 0;
}

static char *slurp(const char *path, char *error, size_t n) {
    FILE *fp = fopen(path, "rb");
    if (!fp) {
        fail(error, n, "Failed to open JSON configuration file", strerror(errno));
        return NULL;
    }
    size_t cap = 4096, len = 0;
    char *buf = (char *)malloc(cap);
    while (buf) {
        const size_t got = fread(buf + len, 1, cap - len - 1, fp);
        len += got;
        if (len + 1 < cap) break;
        cap *= 2;
        char *t = (char *)realloc(buf, cap);
        if (!t) {
            free(buf);
            buf = NULL;
        } else {
            buf = t;
        }
    }
    const int bad = ferror(fp);
    fclose(fp);
    if (!buf || bad) {
        free(buf);
        fail(error, n, "Failed to read JSON configuration file", path);
        return NULL;
    }
    buf[len] = '\0';
    return buf;
}

int ns_config_load_files(NSConfig *c, const char **paths, size_t count, char *error, size_t n) {
    if (!c || (!paths && count > 0)) {
        fail(error, n, "Invalid arguments to ns_config_load_files", NULL);
        return -1;
    }
    for (size_t i = 0; i < count; ++i) {
        const char *path = paths[i];
        const char *dot = path ? strrchr(path, '.') : NULL;
        if (!path) {
            fail(error, n, "Invalid arguments to ns_config_load_files", NULL);
            return -1;
        }
        int rc;
        if (dot && ieq(dot, ".json")) {
            char *text = slurp(path, error, n);
            if (!text) return -1;
            rc = load_json(c, text, error, n);
            free(text);
        } else {
            FILE *fp = fopen(path, "r");
            if (!fp) {
                fail(error, n, "Failed to open INI configuration file", strerror(errno));
                return -1;
            }
            rc = load_ini(c, fp, error, n);
            fclose(fp);
        }
        if (rc != 0) return -1;
    }
    return 0;
}

int ns_config_validate(const NSConfig *c, const NSConfigSchemaEntry *schema, size_t count,
                       char *error, size_t n) {
    if (!c || !schema) {
        fail(error, n, "Invalid arguments to ns_config_validate", NULL);
        return -1;
    }
    for (size_t i = 0; i < count; ++i) {
        const NSConfigSchemaEntry *s = &schema[i];
        const NSConfigEntry *e = find(c, s->key);
        if (!e) {
            if (s->required) {
                fail(error, n, "Missing required configuration key", s->key);
                return -1;
            }
            continue;
        }
        const char *v = e->value;
        switch (s->type) {
        case NS_CONFIG_VALUE_STRING:
            if (s->allowed_values && s->allowed_count > 0) {
                size_t a = 0;
                while (a < s->allowed_count && !ieq(v, s->allowed_values[a])) ++a;
                if (a == s->allowed_count) {
                    fail(error, n, "Configuration string value not permitted", s->key);
                    return -1;
                }
            }
            break;
        case NS_CONFIG_VALUE_INT: {
            long x;
            if (!parse_long(v, &x)) {
                fail(error, n, "Configuration key expects integer value", s->key);
                return -1;
            }
            if (s->has_min && x < s->min_value) {
                fail(error, n, "Configuration integer below minimum", s->key);
                return -1;
            }
            if (s->has_max && x > s->max_value) {
                fail(error, n, "Configuration integer above maximum", s->key);
                return -1;
            }
            break;
        }
        case NS_CONFIG_VALUE_DOUBLE: {
            double x;
            if (!parse_double(v, &x)) {
                fail(error, n, "Configuration key expects numeric value", s->key);
                return -1;
            }
            if (s->has_min && x < s->min_value) {
                fail(error, n, "Configuration value below minimum", s->key);
                return -1;
            }
            if (s->has_max && x > s->max_value) {
                fail(error, n, "Configuration value above maximum", s->key);
                return -1;
            }
            break;
        }
        case NS_CONFIG_VALUE_BOOL:
            if (parse_bool(v) < 0) {
                fail(error, n, "Configuration key expects boolean value", s->key);
                return -1;
            }
            break;
        }
    }
    return 0;
}
