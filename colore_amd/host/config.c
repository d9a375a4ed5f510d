/* config.c -- reader for the subset of the libconfig grammar that CoLoRe
 * param files use (the reference links libconfig, io.c:259-261):
 *   setting  := name (':' | '=') value [';' | ',']
 *   value    := scalar | '[' scalars ']' | '(' values ')' | '{' settings '}'
 *   scalar   := "string" ["string"...] | integer (dec/hex, optional L) |
 *               float | true | false (any case)
 *   comments := '#' / '//' to end of line, C-style block comments
 * Lookups take dotted paths ("global.seed").  Integer settings read as
 * floats and integral floats read as integers (CONFIG_OPTION_AUTOCONVERT). */
#include "clr_host.h"
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

enum { T_INT, T_FLOAT, T_BOOL, T_STRING, T_ARRAY, T_GROUP };

typedef struct node {
  char *name;
  int type;
  long long ival;
  double fval;
  char *sval;
  struct node **kids;
  int nkids, cap;
} node;

struct clr_cfg { node *root; };

typedef struct { const char *s; int pos, line; char *err; int errlen; int bad; } lexer;

static void set_err(lexer *L, const char *msg) {
  if (!L->bad) snprintf(L->err, L->errlen, "line %d: %s", L->line, msg);
  L->bad = 1;
}

static void skip_ws(lexer *L) {
  for (;;) {
    const char c = L->s[L->pos];
    if (c == '\n') { L->line++; L->pos++; }
    else if (isspace((unsigned char)c)) L->pos++;
    else if (c == '#' || (c == '/' && L->s[L->pos + 1] == '/')) {
      while (L->s[L->pos] && L->s[L->pos] != '\n') L->pos++;
    } else if (c == '/' && L->s[L->pos + 1] == '*') {
      L->pos += 2;
      while (L->s[L->pos] && !(L->s[L->pos] == '*' && L->s[L->pos + 1] == '/')) {
        if (L->s[L->pos] == '\n') L->line++;
        L->pos++;
      }
      if (L->s[L->pos]) L->pos += 2;
    } else return;
  }
}

static node *new_node(int type) {
  node *n = calloc(1, sizeof(node));
  n->type = type;
  return n;
}

static void add_kid(node *p, node *k) {
  if (p->nkids == p->cap) {
    p->cap = p->cap ? 2 * p->cap : 8;
    p->kids = realloc(p->kids, p->cap * sizeof(node *));
  }
  p->kids[p->nkids++] = k;
}

static void free_node(node *n) {
  if (!n) return;
  for (int i = 0; i < n->nkids; i++) free_node(n->kids[i]);
  free(n->kids);
  free(n->name);
  free(n->sval);
  free(n);
}

static node *parse_value(lexer *L);

static void parse_settings(lexer *L, node *grp, char close) {
  for (;;) {
    skip_ws(L);
    const char c = L->s[L->pos];
    if (c == close) { if (close) L->pos++; return; }
    if (!c) { if (close) set_err(L, "unexpected end of file"); return; }
    if (!(isalpha((unsigned char)c) || c == '*')) { set_err(L, "expected a setting name"); return; }
    const int start = L->pos;
    while (isalnum((unsigned char)L->s[L->pos]) || L->s[L->pos] == '_' || L->s[L->pos] == '-' ||
           L->s[L->pos] == '*')
      L->pos++;
    char *name = strndup(L->s + start, L->pos - start);
    skip_ws(L);
    if (L->s[L->pos] != ':' && L->s[L->pos] != '=') { free(name); set_err(L, "expected ':' or '='"); return; }
    L->pos++;
    node *v = parse_value(L);
    if (!v) { free(name); return; }
    v->name = name;
    add_kid(grp, v);
    skip_ws(L);
    if (L->s[L->pos] == ';' || L->s[L->pos] == ',') L->pos++;
    if (L->bad) return;
  }
}

static node *parse_value(lexer *L) {
  skip_ws(L);
  const char c = L->s[L->pos];
  if (c == '{') {
    L->pos++;
    node *g = new_node(T_GROUP);
    parse_settings(L, g, '}');
    if (L->bad) { free_node(g); return NULL; }
    return g;
  }
  if (c == '[' || c == '(') {
    const char close = c == '[' ? ']' : ')';
    L->pos++;
    node *a = new_node(T_ARRAY);
    for (;;) {
      skip_ws(L);
      if (L->s[L->pos] == close) { L->pos++; break; }
      node *v = parse_value(L);
      if (!v) { free_node(a); return NULL; }
      add_kid(a, v);
      skip_ws(L);
      if (L->s[L->pos] == ',') L->pos++;
      else if (L->s[L->pos] != close) { set_err(L, "expected ',' in list"); free_node(a); return NULL; }
    }
    return a;
  }
  if (c == '"') {
    node *s = new_node(T_STRING);
    size_t cap = 64, len = 0;
    char *buf = malloc(cap);
    while (L->s[L->pos] == '"') {
      L->pos++;
      while (L->s[L->pos] && L->s[L->pos] != '"') {
        char ch = L->s[L->pos++];
        if (ch == '\\' && L->s[L->pos]) {
          const char e = L->s[L->pos++];
          ch = e == 'n' ? '\n' : e == 't' ? '\t' : e;
        }
        if (len + 2 > cap) { cap *= 2; buf = realloc(buf, cap); }
        buf[len++] = ch;
      }
      if (L->s[L->pos] != '"') { set_err(L, "unterminated string"); free(buf); free_node(s); return NULL; }
      L->pos++;
      skip_ws(L);
    }
    buf[len] = 0;
    s->sval = buf;
    return s;
  }
  /* bare word: bool or number */
  const int start = L->pos;
  while (L->s[L->pos] && !isspace((unsigned char)L->s[L->pos]) && !strchr(";,]){}#", L->s[L->pos]) &&
         !(L->s[L->pos] == '/' && (L->s[L->pos + 1] == '/' || L->s[L->pos + 1] == '*')))
    L->pos++;
  if (L->pos == start) { set_err(L, "expected a value"); return NULL; }
  char *w = strndup(L->s + start, L->pos - start);
  node *n = NULL;
  if (!strcasecmp(w, "true") || !strcasecmp(w, "false")) {
    n = new_node(T_BOOL);
    n->ival = !strcasecmp(w, "true");
  } else {
    char *end;
    size_t wl = strlen(w);
    int is_float = strpbrk(w, ".eE") != NULL && !(wl > 2 && w[0] == '0' && (w[1] == 'x' || w[1] == 'X'));
    if (!is_float && wl > 1 && (w[wl - 1] == 'L' || w[wl - 1] == 'l')) w[--wl] = 0;
    errno = 0;
    if (is_float) {
      const double v = strtod(w, &end);
      if (*end || errno) { free(w); set_err(L, "bad number"); return NULL; }
      n = new_node(T_FLOAT);
      n->fval = v;
    } else {
      const long long v = strtoll(w, &end, 0);
      if (*end || errno) { free(w); set_err(L, "bad value"); return NULL; }
      n = new_node(T_INT);
      n->ival = v;
    }
  }
  free(w);
  return n;
}

clr_cfg *clr_cfg_read(const char *fname, char *err, int errlen) {
  FILE *f = fopen(fname, "rb");
  if (!f) { snprintf(err, errlen, "Couldn't open file %s", fname); return NULL; }
  fseek(f, 0, SEEK_END);
  const long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = malloc(sz + 1);
  if ((long)fread(buf, 1, sz, f) != sz) { fclose(f); free(buf); snprintf(err, errlen, "Couldn't read %s", fname); return NULL; }
  buf[sz] = 0;
  fclose(f);
  lexer L = {buf, 0, 1, err, errlen, 0};
  node *root = new_node(T_GROUP);
  parse_settings(&L, root, 0);
  free(buf);
  if (L.bad) { free_node(root); return NULL; }
  clr_cfg *c = malloc(sizeof(clr_cfg));
  c->root = root;
  return c;
}

void clr_cfg_free(clr_cfg *c) {
  if (!c) return;
  free_node(c->root);
  free(c);
}

static const node *lookup(const clr_cfg *c, const char *path) {
  const node *n = c->root;
  char *p = strdup(path), *save = NULL;
  for (char *tok = strtok_r(p, ".", &save); tok && n; tok = strtok_r(NULL, ".", &save)) {
    const node *next = NULL;
    if (n->type == T_GROUP)
      for (int i = 0; i < n->nkids; i++)
        if (n->kids[i]->name && !strcmp(n->kids[i]->name, tok)) { next = n->kids[i]; break; }
    n = next;
  }
  free(p);
  return n;
}

int clr_cfg_has(const clr_cfg *c, const char *path) { return lookup(c, path) != NULL; }

int clr_cfg_string(const clr_cfg *c, const char *path, char *out, int outlen) {
  const node *n = lookup(c, path);
  if (!n || n->type != T_STRING) return 0;
  snprintf(out, outlen, "%s", n->sval);
  return 1;
}

int clr_cfg_double(const clr_cfg *c, const char *path, double *out) {
  const node *n = lookup(c, path);
  if (!n) return 0;
  if (n->type == T_FLOAT) { *out = n->fval; return 1; }
  if (n->type == T_INT) { *out = (double)n->ival; return 1; }
  return 0;
}

int clr_cfg_int(const clr_cfg *c, const char *path, int *out) {
  const node *n = lookup(c, path);
  if (!n) return 0;
  if (n->type == T_INT) { *out = (int)n->ival; return 1; }
  if (n->type == T_FLOAT && n->fval == floor(n->fval) && fabs(n->fval) < 2147483648.0) { *out = (int)n->fval; return 1; }
  return 0;
}

int clr_cfg_bool(const clr_cfg *c, const char *path, int *out) {
  const node *n = lookup(c, path);
  if (!n || n->type != T_BOOL) return 0;
  *out = (int)n->ival;
  return 1;
}

int clr_cfg_double_array(const clr_cfg *c, const char *path, double *out, int nmax) {
  const node *n = lookup(c, path);
  if (!n || n->type != T_ARRAY) return -1;
  if (n->nkids > nmax) return -2;
  for (int i = 0; i < n->nkids; i++) {
    const node *k = n->kids[i];
    if (k->type == T_FLOAT) out[i] = k->fval;
    else if (k->type == T_INT) out[i] = (double)k->ival;
    else return -1;
  }
  return n->nkids;
}
