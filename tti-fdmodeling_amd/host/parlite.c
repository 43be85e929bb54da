/* parlite.c -- command-line parameters and diagnostics (see parlite.h). */
#include "parlite.h"

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  char *key, *val;
} kv_t;

static kv_t *g_kv = NULL;
static int g_nkv = 0, g_cap = 0;
static int g_argc = 0;
static char *g_prog = "fdmod";
static const char **g_sdoc = NULL;

static void add_kv(const char *word)
{
  const char *eq = strchr(word, '=');
  if (!eq || eq == word) return;
  if (g_nkv == g_cap) {
    g_cap = g_cap ? 2 * g_cap : 32;
    g_kv = (kv_t *)realloc(g_kv, sizeof(kv_t) * g_cap);
  }
  size_t kl = (size_t)(eq - word);
  g_kv[g_nkv].key = (char *)malloc(kl + 1);
  memcpy(g_kv[g_nkv].key, word, kl);
  g_kv[g_nkv].key[kl] = 0;
  g_kv[g_nkv].val = strdup(eq + 1);
  g_nkv++;
}

static void read_parfile(const char *path)
{
  FILE *fp = fopen(path, "r");
  if (!fp) err("cannot open parameter file %s", path);
  char line[4096];
  while (fgets(line, sizeof line, fp)) {
    char *h = strchr(line, '#');
    if (h) *h = 0;
    char *tok = strtok(line, " \t\r\n");
    while (tok) {
      add_kv(tok);
      tok = strtok(NULL, " \t\r\n");
    }
  }
  fclose(fp);
}

void initargs(int argc, char **argv)
{
  g_argc = argc;
  if (argc > 0 && argv[0]) {
    const char *s = strrchr(argv[0], '/');
    g_prog = strdup(s ? s + 1 : argv[0]);
  }
  for (int i = 1; i < argc; i++) {
    if (strncmp(argv[i], "par=", 4) == 0) read_parfile(argv[i] + 4);
    else add_kv(argv[i]);
  }
}

void set_selfdoc(const char **sdoc) { g_sdoc = sdoc; }

void requestdoc(int flag)
{
  (void)flag;
  if (g_argc > 1) return;
  if (g_sdoc)
    for (int i = 0; g_sdoc[i]; i++) fprintf(stderr, "%s\n", g_sdoc[i]);
  exit(1);
}

static const char *lookup(const char *name)
{
  for (int i = g_nkv - 1; i >= 0; i--)
    if (strcmp(g_kv[i].key, name) == 0) return g_kv[i].val;
  return NULL;
}

int getparint(const char *name, int *v)
{
  const char *s = lookup(name);
  if (!s) return 0;
  char *end;
  long x = strtol(s, &end, 10);
  if (end == s) err("parameter %s=%s is not an integer", name, s);
  *v = (int)x;
  return 1;
}

int getparfloat(const char *name, float *v)
{
  const char *s = lookup(name);
  if (!s) return 0;
  char *end;
  double x = strtod(s, &end);
  if (end == s) err("parameter %s=%s is not a number", name, s);
  *v = (float)x;
  return 1;
}

int getparstring(const char *name, char **v)
{
  const char *s = lookup(name);
  if (!s) return 0;
  *v = (char *)s;
  return 1;
}

const char *progname(void) { return g_prog; }

void err(const char *fmt, ...)
{
  va_list ap;
  fflush(stdout);
  fprintf(stderr, "\n%s: ", g_prog);
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
  exit(1);
}

void warn(const char *fmt, ...)
{
  va_list ap;
  fflush(stdout);
  fprintf(stderr, "\n%s: ", g_prog);
  va_start(ap, fmt);
  vfprintf(stderr, fmt, ap);
  va_end(ap);
  fprintf(stderr, "\n");
}
