/*
 * NxStereo.HipSegment.Nif: the Erlang NIF over include/nxs_hip.h that NxStereo.HipSegment
 * runs its segments through, in place of the stereo_match Port behind NxStereo.Matcher
 * (lib/nx_stereo/matcher.ex:248-255). Compiled where the Erlang runtime's headers are:
 *
 *   cc -O2 -fPIC -shared -I$ERTS_INCLUDE_DIR -I../../../include nxs_nif.c \
 *      -L../.. -lnxs_hip -Wl,-rpath,'$ORIGIN' -o priv/nxs_nif.so
 *
 * open(device)          -> {:ok, ctx} | {:error, reason}: an nxs_ctx (the HIP streams and
 *                          device memory of every segment configured through it), closed
 *                          when the resource is collected or by close/1.
 * submit(ctx, record)   -> {:ok, ref} | {:error, reason}: queues one protocol.rs record
 *                          (configure 1, run 2, release 3) and returns at once; its reply
 *                          record arrives as {:nxs_reply, ref, binary} at the calling
 *                          process, posted from the context's completion thread. Replies
 *                          of one segment come in submission order.
 * close(ctx)            -> :ok: fails the requests not started, waits for the runs on the
 *                          device, releases every segment.
 *
 * No call blocks on the GPU, so no dirty scheduler is needed.
 */
#include <erl_nif.h>
#include <string.h>

#include "nxs_hip.h"

static ErlNifResourceType *CONTEXT;

typedef struct {
    ErlNifMutex *lock;
    nxs_ctx *ctx;
} context_resource;

/* What a reply goes back with: the submitting process and the request's reference. */
typedef struct {
    ErlNifPid owner;
    ErlNifEnv *env;
    ERL_NIF_TERM ref;
} waiter;

static ERL_NIF_TERM error_tuple(ErlNifEnv *env, const char *reason) {
    ERL_NIF_TERM binary;
    size_t n = strlen(reason);
    memcpy(enif_make_new_binary(env, n, &binary), reason, n);
    return enif_make_tuple2(env, enif_make_atom(env, "error"), binary);
}

static void close_context(context_resource *r) {
    enif_mutex_lock(r->lock);
    nxs_ctx *ctx = r->ctx;
    r->ctx = NULL;
    enif_mutex_unlock(r->lock);
    if (ctx) nxs_close(ctx);
}

static void context_dtor(ErlNifEnv *env, void *object) {
    context_resource *r = (context_resource *)object;
    close_context(r);
    enif_mutex_destroy(r->lock);
}

/* The context's completion thread: the reply copied into a binary, sent to the waiter. */
static void done(void *user, int64_t ticket, const uint8_t *reply, size_t len) {
    waiter *w = (waiter *)user;
    ErlNifEnv *env = enif_alloc_env();
    ERL_NIF_TERM binary;
    if (len) memcpy(enif_make_new_binary(env, len, &binary), reply, len);
    else enif_make_new_binary(env, 0, &binary);
    enif_send(NULL, &w->owner, env,
              enif_make_tuple3(env, enif_make_atom(env, "nxs_reply"), enif_make_copy(env, w->ref),
                               binary));
    enif_free_env(env);
    enif_free_env(w->env);
    enif_free(w);
    (void)ticket;
}

static ERL_NIF_TERM open_nif(ErlNifEnv *env, int argc, const ERL_NIF_TERM argv[]) {
    int device;
    char reason[512];
    if (argc != 1 || !enif_get_int(env, argv[0], &device)) return enif_make_badarg(env);
    context_resource *r = enif_alloc_resource(CONTEXT, sizeof *r);
    r->lock = enif_mutex_create("nxs_ctx");
    r->ctx = NULL;
    if (nxs_open(device, &r->ctx, reason, sizeof reason) != NXS_OK) {
        r->ctx = NULL;
        enif_release_resource(r);
        return error_tuple(env, reason);
    }
    ERL_NIF_TERM term = enif_make_resource(env, r);
    enif_release_resource(r);
    return enif_make_tuple2(env, enif_make_atom(env, "ok"), term);
}

static ERL_NIF_TERM submit_nif(ErlNifEnv *env, int argc, const ERL_NIF_TERM argv[]) {
    context_resource *r;
    ErlNifBinary record;
    if (argc != 2 || !enif_get_resource(env, argv[0], CONTEXT, (void **)&r) ||
        !enif_inspect_iolist_as_binary(env, argv[1], &record))
        return enif_make_badarg(env);
    waiter *w = enif_alloc(sizeof *w);
    enif_self(env, &w->owner);
    w->env = enif_alloc_env();
    w->ref = enif_make_ref(w->env);
    ERL_NIF_TERM ref = enif_make_copy(env, w->ref);
    enif_mutex_lock(r->lock);
    /* nxs_submit copies the record before it returns. */
    int64_t ticket = r->ctx ? nxs_submit(r->ctx, record.data, record.size, done, w) : -1;
    enif_mutex_unlock(r->lock);
    if (ticket < 0) {
        enif_free_env(w->env);
        enif_free(w);
        return error_tuple(env, r->ctx ? nxs_last_error() : "the context is closed");
    }
    return enif_make_tuple2(env, enif_make_atom(env, "ok"), ref);
}

static ERL_NIF_TERM close_nif(ErlNifEnv *env, int argc, const ERL_NIF_TERM argv[]) {
    context_resource *r;
    if (argc != 1 || !enif_get_resource(env, argv[0], CONTEXT, (void **)&r))
        return enif_make_badarg(env);
    close_context(r);
    return enif_make_atom(env, "ok");
}

static int load(ErlNifEnv *env, void **priv, ERL_NIF_TERM info) {
    CONTEXT = enif_open_resource_type(env, NULL, "nxs_ctx", context_dtor, ERL_NIF_RT_CREATE, NULL);
    return CONTEXT ? 0 : 1;
}

static ErlNifFunc functions[] = {
    {"open", 1, open_nif, 0},
    {"submit", 2, submit_nif, 0},
    /* nxs_close waits for the runs on the device: a dirty I/O scheduler. */
    {"close", 1, close_nif, ERL_NIF_DIRTY_JOB_IO_BOUND},
};

ERL_NIF_INIT(Elixir.NxStereo.HipSegment.Nif, functions, load, NULL, NULL, NULL)
