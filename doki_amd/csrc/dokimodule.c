/*
 * dokimodule.c — the `doki` CPython extension of the MI355X state-vector
 * core.  Same module name, method table, positional signatures, capsule
 * names, exception (`qsimov.doki.error`) and error texts as the reference's
 * extension (Mowstyl/Doki 1.6.1, src/doki/doki.c:114-222), so callers switch
 * without changes.  All state work goes through the C ABI of libdoki_hip.so
 * (include/doki_hip.h); there is no CPU compute path for states here.
 *
 * Deliberate differences (stricter, never silently different):
 *  - registry_get / registry_prob / funmatrix ids are bounds-checked and
 *    raise doki.error instead of reading out of bounds (reference doki.c:570,
 *    :1227 do not check);
 *  - repeated target qubits raise doki.error (doki.c:866-888 does not check);
 *  - a registry passed to registry_del can no longer be used (raises instead
 *    of dangling, doki.c:400-415);
 *  - num_threads is validated and ignored (work runs on the GPU);
 *  - funmatrix_create keeps the reference's layout (element (i, j) of an
 *    r x c input at i * r + j) but drops values that would land past the
 *    array (r > c; the reference writes them out of bounds), and reads
 *    there give NaN; funmatrix_mem of a node over a created matrix returns
 *    its byte count (the reference raises SystemError there);
 *  - constructors of more than 62 qubits raise doki.error (2^n overflows).
 */
#define PY_SSIZE_T_CLEAN
#include <Python.h>
#include <complex.h>
#include <math.h>
#include <numpy/arrayobject.h>
#include <stdint.h>
#include <string.h>

#include "doki_hip.h"

/* not in the public header: library teardown guard */
extern void dk_shutdown(void);

#define CAP_STATE "qsimov.doki.state_vector"
#define CAP_GATE "qsimov.doki.gate"
#define CAP_FMAT "qsimov.doki.funmatrix"

/* sizeof(struct FMatrix) and sizeof(struct DMatrixForTrace) of the reference
 * on LP64 (funmatrix.h:28-66, funmatrix.c:31-37); funmatrix_mem reports them
 * (funmatrix.c:1550-1567). */
#define REF_FMATRIX_BYTES 112
#define REF_PTRACE_ARGV_BYTES 24

typedef double _Complex cplx;

static PyObject *DokiError;

static PyObject *dk_fail(int rc, const char *fallback)
{
	const char *m = dk_last_error();
	(void)rc;
	PyErr_SetString(DokiError, (m && m[0]) ? m : fallback);
	return NULL;
}

static int check_threads(int num_threads)
{
	if (num_threads <= 0 && num_threads != -1) {
		PyErr_SetString(DokiError, "num_threads must be at least 1 (or -1 "
					   "to let OpenMP choose)");
		return -1;
	}
	return 0;
}

/* A gate capsule: its matrix on the host, and on the device once used. */
typedef struct {
	unsigned k;
	unsigned dim;
	double *mat;     /* dim*dim interleaved, row-major */
	dk_gate *dev;    /* created on first use */
} gate_t;

static dk_gate *gate_device(gate_t *g)
{
	if (!g->dev) {
		int rc = dk_gate_new(g->k, g->mat, &g->dev);
		if (rc) {
			dk_fail(rc, "Failed to apply gate");
			return NULL;
		}
	}
	return g->dev;
}

/* ------------------------------------------------------------------------ */
/* registries                                                               */
/* ------------------------------------------------------------------------ */

#define CHUNK 4096 /* amplitudes per read-back chunk of registry_get */

/*
 * Registries are values (every registry_apply returns a new one, doki.c:939,
 * qops.c:192), so the device work behind them may run later than the call
 * that asked for it, as long as every read sees the reference's value.  A
 * registry is either materialised (st holds its state) or pending: it is
 * its parent registry with one more gate.  registry_apply makes pending
 * registries; anything that reads a registry materialises it first, running
 * the chain of pending gates from the nearest materialised ancestor:
 *
 *  - EXEC_FUSE (the default): the whole chain at once, as cache-blocked
 *    passes (dk_circuit_passes: one state pass for many gates);
 *  - EXEC_GATES: one gate per registry_apply, deferred by one call (the call
 *    that applies a gate to a pending registry first runs that registry's
 *    own gate), so each gate runs when it is known whether its input is
 *    still referenced;
 *  - EXEC_EAGER: each registry_apply launches its gate at once, out of place,
 *    as the reference does.
 *
 * A chain runs in place on its ancestor's state when nothing else can read
 * that ancestor: its capsule is referenced only by the chain (Python
 * reference count 1, or the registry was deleted) and no other pending
 * registry derives from it.  Otherwise the first gate writes a new state.
 * Pending registries that something else references (a caller's variable,
 * another pending registry, a density) are materialised on the way, so a
 * branch taken from an intermediate registry reads that registry's value.
 */
enum { EXEC_EAGER = 0, EXEC_GATES = 1, EXEC_FUSE = 2 };
static int g_exec = EXEC_FUSE;
/* pending gates after which registry_apply materialises a chain anyway */
#define FUSE_MAX_PENDING 2048

typedef struct {
	dk_state *st;       /* the state; NULL while pending or once deleted */
	unsigned nq;
	int deleted;        /* registry_del was called */
	/* pending: this registry is `parent` with `gate` applied */
	PyObject *parent;   /* registry capsule (a reference is held) */
	PyObject *gate;     /* gate capsule (a reference is held) */
	unsigned targets[DK_MAX_GATE_QUBITS];
	uint64_t cm, am;
	unsigned depth;     /* pending gates from the materialised ancestor */
	unsigned children;  /* pending registries whose parent this is */
	int norm_known;
	double norm;
	int phase_known;
	double phase;
	uint64_t cache_first; /* read-back chunk of raw amplitudes */
	uint64_t cache_count;
	double *cache;
	double *probs;       /* all single-qubit probabilities, or NULL */
} reg_t;

static reg_t *reg_of(PyObject *cap)
{
	return (reg_t *)PyCapsule_GetPointer(cap, CAP_STATE);
}

/* Detaches a pending registry from its parent (the caller releases the
 * returned reference).  A deleted parent that no pending registry needs any
 * more gives its state back at once. */
static PyObject *reg_detach(reg_t *r)
{
	PyObject *p = r->parent;
	r->parent = NULL;
	Py_CLEAR(r->gate);
	if (p) {
		reg_t *pr = reg_of(p);
		if (pr) {
			pr->children--;
			if (pr->deleted && pr->children == 0 && pr->st) {
				dk_state_free(pr->st);
				pr->st = NULL;
			}
		}
	}
	return p;
}

/* Drops a reference to registry capsule p and, while that was the last
 * one, to its pending ancestors in turn: each is detached from its parent
 * before it is destroyed, so a long chain of pending registries goes
 * without recursion. */
static void reg_release_chain(PyObject *p)
{
	while (p) {
		PyObject *pp = NULL;
		if (Py_REFCNT(p) == 1) {
			reg_t *pr = reg_of(p);
			if (pr)
				pp = reg_detach(pr);
			else
				PyErr_Clear();
		}
		Py_DECREF(p);
		p = pp;
	}
}

static void reg_free(reg_t *r)
{
	if (!r)
		return;
	reg_release_chain(reg_detach(r));
	if (r->st)
		dk_state_free(r->st);
	PyMem_Free(r->cache);
	PyMem_Free(r->probs);
	PyMem_Free(r);
}

static void reg_capsule_destroy(PyObject *cap)
{
	reg_t *r = reg_of(cap);
	if (r)
		reg_free(r);
}

static PyObject *reg_capsule(reg_t *r)
{
	PyObject *cap = PyCapsule_New(r, CAP_STATE, reg_capsule_destroy);
	if (!cap)
		reg_free(r);
	return cap;
}

static PyObject *reg_wrap(dk_state *st)
{
	reg_t *r = (reg_t *)PyMem_Calloc(1, sizeof(reg_t));
	if (!r) {
		dk_state_free(st);
		return PyErr_NoMemory();
	}
	r->st = st;
	r->nq = dk_state_num_qubits(st);
	return reg_capsule(r);
}

/* The registry of a capsule, materialised or not, or NULL with doki.error
 * set. */
static reg_t *reg_peek(PyObject *cap, const char *nullmsg)
{
	reg_t *r = reg_of(cap);
	if (!r) {
		PyErr_SetString(DokiError, nullmsg);
		return NULL;
	}
	if (r->deleted) {
		PyErr_SetString(DokiError, "registry has been deleted");
		return NULL;
	}
	return r;
}

/* Runs `count` pending gates (nodes[0] first) on `in`, in place or into a
 * new state (*out). */
static int run_gates(dk_state *in, int in_place, reg_t **nodes, unsigned count,
		     dk_state **out)
{
	dk_circuit *c = NULL;
	int rc;
	*out = NULL;
	if (count > 1 && g_exec == EXEC_FUSE) {
		rc = dk_circuit_new(dk_state_num_qubits(in), &c);
		for (unsigned i = 0; !rc && i < count; i++) {
			gate_t *g = (gate_t *)PyCapsule_GetPointer(nodes[i]->gate, CAP_GATE);
			rc = dk_circuit_add(c, g->k, g->mat, nodes[i]->targets,
					    nodes[i]->cm, nodes[i]->am);
		}
		if (!rc)
			rc = dk_circuit_passes(c, NULL);
		if (!rc)
			rc = dk_circuit_run(c, in, in_place, out);
		dk_circuit_free(c);
		return rc;
	}
	for (unsigned i = 0; i < count; i++) {
		gate_t *g = (gate_t *)PyCapsule_GetPointer(nodes[i]->gate, CAP_GATE);
		dk_gate *dg = gate_device(g);
		dk_state *o = NULL;
		if (!dg)
			return DK_E_ALLOC;
		rc = dk_apply(i == 0 ? in : *out, dg, nodes[i]->targets, nodes[i]->cm,
			      nodes[i]->am, i > 0 || in_place, &o);
		if (rc) {
			if (i > 0 && !in_place)
				dk_state_free(*out);
			*out = NULL;
			return rc;
		}
		*out = o;
	}
	return DK_OK;
}

/* Materialises a pending registry (see EXEC_FUSE above); 0, or -1 with an
 * exception set. */
static int reg_materialise(PyObject *cap)
{
	reg_t *n = reg_of(cap);
	PyObject **path;
	reg_t **nodes;
	unsigned len = 0, cap_len, i, root, seg;
	int rc = 0;
	if (n->st)
		return 0;
	cap_len = n->depth + 1;
	path = (PyObject **)PyMem_Malloc(cap_len * sizeof(PyObject *));
	nodes = (reg_t **)PyMem_Malloc(cap_len * sizeof(reg_t *));
	if (!path || !nodes) {
		PyMem_Free(path);
		PyMem_Free(nodes);
		PyErr_NoMemory();
		return -1;
	}
	/* path[0] = this registry, ..., path[root] = materialised ancestor */
	path[len++] = cap;
	while (!reg_of(path[len - 1])->st) {
		PyObject *p = reg_of(path[len - 1])->parent;
		if (!p || len == cap_len) {
			PyMem_Free(path);
			PyMem_Free(nodes);
			PyErr_SetString(DokiError, "registry has lost its source state");
			return -1;
		}
		path[len++] = p;
	}
	root = len - 1;
	/* segments from the root down: each ends at this registry or at a
	 * pending registry something else references, which keeps its value */
	seg = root;
	for (i = root; i-- > 0;) {
		reg_t *x = reg_of(path[i]);
		if (i > 0 && Py_REFCNT(path[i]) == 1)
			continue;
		{
			reg_t *src = reg_of(path[seg]);
			unsigned cnt = seg - i, j;
			int in_place = src->children == 1 &&
				       (Py_REFCNT(path[seg]) == 1 || src->deleted);
			dk_state *in = src->st, *out = NULL;
			for (j = 0; j < cnt; j++)
				nodes[j] = reg_of(path[seg - 1 - j]);
			if (in_place)
				src->st = NULL;     /* the chain takes the state */
			rc = run_gates(in, in_place, nodes, cnt, &out);
			if (rc) {
				if (in_place)
					src->st = in;
				dk_fail(rc, "Failed to apply gate");
				break;
			}
			x->st = out;
			seg = i;
		}
	}
	if (!rc) {
		/* every registry on the path below the root now holds or no
		 * longer needs its source: release the links (no cascade) */
		for (i = 0; i < root; i++)
			path[i] = reg_detach(reg_of(path[i]));
		for (i = 0; i < root; i++)
			Py_XDECREF(path[i]);
	}
	PyMem_Free(path);
	PyMem_Free(nodes);
	return rc ? -1 : 0;
}

/* The registry of a capsule with its state materialised, or NULL with
 * doki.error set. */
static reg_t *reg_get(PyObject *cap, const char *nullmsg)
{
	reg_t *r = reg_peek(cap, nullmsg);
	if (!r)
		return NULL;
	if (!r->st && reg_materialise(cap))
		return NULL;
	return r;
}

static int reg_norm(reg_t *r)
{
	if (!r->norm_known) {
		int rc = dk_state_norm(r->st, &r->norm);
		if (rc) {
			dk_fail(rc, "Failed to read state");
			return -1;
		}
		r->norm_known = 1;
	}
	return 0;
}

/* Normalised amplitude id: raw / norm with the C99 complex-by-real division
 * of state_get (qstate.h:86). */
static int reg_value(reg_t *r, uint64_t id, cplx *out)
{
	uint64_t len = (uint64_t)1 << r->nq;
	if (reg_norm(r))
		return -1;
	if (!r->cache || id < r->cache_first ||
	    id >= r->cache_first + r->cache_count) {
		uint64_t first = id - id % CHUNK;
		uint64_t count = len - first < CHUNK ? len - first : CHUNK;
		int rc;
		if (!r->cache) {
			r->cache = (double *)PyMem_Malloc(2 * CHUNK * sizeof(double));
			if (!r->cache) {
				PyErr_NoMemory();
				return -1;
			}
		}
		rc = dk_read_raw(r->st, first, count, r->cache, NULL);
		if (rc) {
			r->cache_count = 0;
			dk_fail(rc, "Failed to read state");
			return -1;
		}
		r->cache_first = first;
		r->cache_count = count;
	}
	{
		uint64_t k = id - r->cache_first;
		cplx raw = r->cache[2 * k] + I * r->cache[2 * k + 1];
		*out = raw / r->norm;
	}
	return 0;
}

/* get_global_phase, qops.c:46-70: arg of the first nonzero normalised
 * amplitude when its imaginary part is nonzero, else 0; cached. */
static int reg_phase(reg_t *r, double *phase)
{
	if (!r->phase_known) {
		uint64_t idx = 0, len = (uint64_t)1 << r->nq;
		double ph = 0.0;
		int rc = dk_first_nonzero(r->st, &idx);
		if (rc) {
			dk_fail(rc, "Failed to read state");
			return -1;
		}
		if (idx < len) {
			cplx v;
			if (reg_value(r, idx, &v))
				return -1;
			if (cimag(v) != 0.)
				ph = carg(v);
		}
		r->phase = ph;
		r->phase_known = 1;
	}
	*phase = r->phase;
	return 0;
}

static PyObject *doki_registry_new(PyObject *self, PyObject *args)
{
	unsigned int nq;
	int verbose;
	dk_state *st;
	int rc;
	(void)self;
	if (!PyArg_ParseTuple(args, "Ip", &nq, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_new(num_qubits, verbose)");
		return NULL;
	}
	if (nq == 0) {
		PyErr_SetString(DokiError, "num_qubits can't be zero");
		return NULL;
	}
	if (nq > DK_MAX_QUBITS) {
		PyErr_SetString(DokiError, "Number of qubits exceeds maximum");
		return NULL;
	}
	rc = dk_state_new(nq, 1, &st);
	if (rc)
		return dk_fail(rc, "Unknown error when creating state");
	return reg_wrap(st);
}

static PyObject *doki_registry_clone(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int nthreads, verbose, rc;
	reg_t *r;
	dk_state *st;
	(void)self;
	if (!PyArg_ParseTuple(args, "Oip", &cap, &nthreads, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: registry_clone(registry, num_threads, verbose)");
		return NULL;
	}
	if (check_threads(nthreads))
		return NULL;
	r = reg_get(cap, "NULL pointer to source registry");
	if (!r)
		return NULL;
	rc = dk_state_clone(r->st, &st);
	if (rc)
		return dk_fail(rc, "Unknown error when cloning state");
	return reg_wrap(st);
}

static PyObject *doki_registry_del(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	reg_t *r;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_del(registry, verbose)");
		return NULL;
	}
	r = reg_of(cap);
	if (r) {
		/* freed at once unless pending registries still derive from it
		 * (then when the last of them no longer needs it) */
		r->deleted = 1;
		if (r->children == 0) {
			if (r->st)
				dk_state_free(r->st);
			r->st = NULL;
			reg_release_chain(reg_detach(r));
		}
		PyMem_Free(r->cache);
		r->cache = NULL;
		r->cache_count = 0;
	} else {
		PyErr_Clear();
	}
	Py_RETURN_NONE;
}

static int parse_complex(PyObject *o, cplx *out)
{
	if (PyComplex_Check(o)) {
		*out = PyComplex_RealAsDouble(o) + I * PyComplex_ImagAsDouble(o);
	} else if (PyFloat_Check(o)) {
		*out = PyFloat_AsDouble(o) + I * 0.0;
	} else if (PyLong_Check(o)) {
		*out = (double)PyLong_AsLong(o) + I * 0.0;
	} else {
		return -1;
	}
	return 0;
}

static PyObject *doki_registry_new_data(PyObject *self, PyObject *args)
{
	PyObject *vals;
	unsigned int nq;
	short verbose;
	uint64_t len, i;
	double *buf = NULL;
	dk_state *st;
	int rc;
	(void)self;
	if (!PyArg_ParseTuple(args, "IOh", &nq, &vals, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: registry_new_data(num_qubits, values, verbose)");
		return NULL;
	}
	if (nq == 0) {
		PyErr_SetString(DokiError, "num_qubits can't be zero");
		return NULL;
	}
	if (nq > DK_MAX_QUBITS) {
		PyErr_SetString(DokiError, "Number of qubits exceeds maximum");
		return NULL;
	}
	len = (uint64_t)1 << nq;
	if (PyArray_Check(vals)) {
		PyArrayObject *arr = (PyArrayObject *)vals, *c;
		if (!PyArray_ISNUMBER(arr)) {
			PyErr_SetString(DokiError, "values have to be numbers");
			return NULL;
		}
		if ((uint64_t)PyArray_SIZE(arr) != len) {
			PyErr_SetString(DokiError, "Wrong array size for the specified "
						   "number of qubits");
			return NULL;
		}
		c = (PyArrayObject *)PyArray_FROMANY(vals, NPY_CDOUBLE, 0, 0,
						     NPY_ARRAY_C_CONTIGUOUS |
							     NPY_ARRAY_ALIGNED);
		if (!c)
			return NULL;
		rc = dk_state_from_host(nq, (const double *)PyArray_DATA(c), &st);
		Py_DECREF(c);
	} else if (PyList_Check(vals)) {
		if ((uint64_t)PyList_GET_SIZE(vals) != len) {
			PyErr_SetString(DokiError, "Wrong list size for the specified "
						   "number of qubits\n");
			return NULL;
		}
		buf = (double *)PyMem_Malloc(2 * len * sizeof(double));
		if (!buf)
			return PyErr_NoMemory();
		for (i = 0; i < len; i++) {
			PyObject *o = PyList_GET_ITEM(vals, i);
			double re = PyComplex_RealAsDouble(o);
			double im = PyComplex_ImagAsDouble(o);
			if (PyErr_Occurred()) {
				PyMem_Free(buf);
				PyErr_Clear();
				PyErr_SetString(DokiError, "values have to be numbers");
				return NULL;
			}
			buf[2 * i] = re;
			buf[2 * i + 1] = im;
		}
		rc = dk_state_from_host(nq, buf, &st);
		PyMem_Free(buf);
	} else {
		PyErr_SetString(DokiError, "values has to be either a python list "
					   "or a numpy array");
		return NULL;
	}
	if (rc)
		return dk_fail(rc, "Unknown error when creating state");
	return reg_wrap(st);
}

/*
 * DLPack interchange of registry values with other libraries on the device
 * (PyTorch, CuPy, ...), without a host round trip.  The structures follow
 * the public DLPack ABI (dlpack.h, version 0.8): a "dltensor" capsule holds
 * a DLManagedTensor whose deleter the consumer calls when it is done; a
 * consumer that takes the tensor renames the capsule "used_dltensor".
 */
typedef struct {
	int32_t device_type;   /* DLDeviceType: 1 CPU, 3 CUDA host, 10 ROCm, 11 ROCm host */
	int32_t device_id;
} dl_device;
typedef struct {
	uint8_t code;          /* 2 float, 5 complex */
	uint8_t bits;
	uint16_t lanes;
} dl_dtype;
typedef struct {
	void *data;
	dl_device device;
	int32_t ndim;
	dl_dtype dtype;
	int64_t *shape;
	int64_t *strides;      /* in elements; NULL for a compact row-major tensor */
	uint64_t byte_offset;
} dl_tensor;
typedef struct dl_managed {
	dl_tensor dl;
	void *manager_ctx;
	void (*deleter)(struct dl_managed *self);
} dl_managed;
enum { DL_CPU = 1, DL_CUDA_HOST = 3, DL_ROCM = 10, DL_ROCM_HOST = 11 };
enum { DL_FLOAT = 2, DL_COMPLEX = 5 };

/* What a produced tensor owns: the state holding the values and the shape. */
typedef struct {
	dl_managed m;
	dk_state *st;
	int64_t shape[1];
} dl_export;

static void dl_export_delete(dl_managed *m)
{
	dl_export *e = (dl_export *)m;
	if (e->st)
		dk_state_free(e->st);
	PyMem_RawFree(e);
}

static void dl_capsule_destroy(PyObject *cap)
{
	/* a capsule nobody consumed still owns its tensor */
	if (PyCapsule_IsValid(cap, "dltensor")) {
		dl_managed *m = (dl_managed *)PyCapsule_GetPointer(cap, "dltensor");
		if (m && m->deleter)
			m->deleter(m);
	}
}

/* registry_to_dlpack(registry, verbose): a "dltensor" capsule with the
 * registry's values (as registry_get returns them, bit for bit) as a 1-D
 * complex128 tensor of 2^n elements in device memory, for
 * torch.utils.dlpack.from_dlpack and the like.  The tensor holds a copy made
 * on the device (registries are values: writes to the tensor do not change
 * the registry). */
static PyObject *doki_registry_to_dlpack(PyObject *self, PyObject *args)
{
	PyObject *cap, *out;
	int verbose, rc, device = 0;
	reg_t *r;
	dl_export *e;
	void *amp = NULL;
	dk_state *vals = NULL;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_to_dlpack(registry, verbose)");
		return NULL;
	}
	r = reg_get(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	rc = dk_state_values(r->st, &vals);
	if (!rc)
		rc = dk_state_device_address(vals, &amp, &device);
	if (rc) {
		if (vals)
			dk_state_free(vals);
		return dk_fail(rc, "Failed to export registry");
	}
	e = (dl_export *)PyMem_RawCalloc(1, sizeof(dl_export));
	if (!e) {
		dk_state_free(vals);
		return PyErr_NoMemory();
	}
	e->st = vals;
	e->shape[0] = (int64_t)1 << r->nq;
	e->m.dl.data = amp;
	e->m.dl.device.device_type = DL_ROCM;
	e->m.dl.device.device_id = device;
	e->m.dl.ndim = 1;
	e->m.dl.dtype.code = DL_COMPLEX;
	e->m.dl.dtype.bits = 128;
	e->m.dl.dtype.lanes = 1;
	e->m.dl.shape = e->shape;
	e->m.dl.strides = NULL;
	e->m.dl.byte_offset = 0;
	e->m.manager_ctx = e;
	e->m.deleter = dl_export_delete;
	out = PyCapsule_New(&e->m, "dltensor", dl_capsule_destroy);
	if (!out)
		dl_export_delete(&e->m);
	return out;
}

/* Whether t is a compact row-major tensor. */
static int dl_compact(const dl_tensor *t)
{
	int64_t want = 1;
	if (!t->strides)
		return 1;
	for (int32_t d = t->ndim - 1; d >= 0; d--) {
		if (t->shape[d] != 1 && t->strides[d] != want)
			return 0;
		want *= t->shape[d];
	}
	return 1;
}

/* registry_from_dlpack(tensor, verbose): a registry holding the values of a
 * DLPack tensor (a "dltensor" capsule, or an object with __dlpack__): 2^n
 * complex128 elements (1-D), or 2^n x 2 float64 (re, im), compact, in
 * device memory of this library's device or in host memory.  Values are
 * taken as given, not normalised, as registry_new_data takes them
 * (doki.c:625-720); the copy stays on the device for device tensors. */
static PyObject *doki_registry_from_dlpack(PyObject *self, PyObject *args)
{
	PyObject *obj, *cap = NULL;
	int verbose, rc, mydev = 0;
	dl_managed *m;
	const dl_tensor *t;
	uint64_t len;
	unsigned nq = 0;
	dk_state *st = NULL;
	const char *why = NULL;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &obj, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_from_dlpack(tensor, verbose)");
		return NULL;
	}
	if (PyCapsule_CheckExact(obj)) {
		cap = obj;
		Py_INCREF(cap);
	} else if (PyObject_HasAttrString(obj, "__dlpack__")) {
		cap = PyObject_CallMethod(obj, "__dlpack__", NULL);
		if (!cap)
			return NULL;
	} else {
		PyErr_SetString(DokiError, "tensor has to be a DLPack capsule or "
					   "provide __dlpack__");
		return NULL;
	}
	if (!PyCapsule_IsValid(cap, "dltensor")) {
		Py_DECREF(cap);
		PyErr_SetString(DokiError, "not an unconsumed DLPack capsule");
		return NULL;
	}
	m = (dl_managed *)PyCapsule_GetPointer(cap, "dltensor");
	t = &m->dl;
	len = 1;
	for (int32_t d = 0; d < t->ndim; d++)
		len *= (uint64_t)t->shape[d];
	if (t->dtype.lanes != 1)
		why = "vector element types are not supported";
	else if (t->dtype.code == DL_COMPLEX && t->dtype.bits == 128 && t->ndim == 1)
		;
	else if (t->dtype.code == DL_FLOAT && t->dtype.bits == 64 && t->ndim == 2 &&
		 t->shape[1] == 2)
		len /= 2;
	else
		why = "values have to be complex128 (n) or float64 (n, 2)";
	if (!why && !dl_compact(t))
		why = "the tensor has to be contiguous";
	if (!why) {
		while (nq < 63 && ((uint64_t)1 << nq) < len)
			nq++;
		if (len < 2 || ((uint64_t)1 << nq) != len || nq > DK_MAX_QUBITS)
			why = "the tensor's length has to be 2^n, 1 <= n <= 62";
	}
	if (!why) {
		const char *data = (const char *)t->data + t->byte_offset;
		if (t->device.device_type == DL_ROCM) {
			rc = dk_device_ordinal(&mydev);
			if (!rc && t->device.device_id != mydev)
				why = "the tensor is on another device";
			else if (!rc)
				rc = dk_state_from_device(nq, data, &st);
		} else if (t->device.device_type == DL_CPU ||
			   t->device.device_type == DL_CUDA_HOST ||
			   t->device.device_type == DL_ROCM_HOST) {
			rc = dk_state_from_host(nq, (const double *)data, &st);
		} else {
			why = "the tensor has to be in host memory or on this device";
		}
		if (!why && rc) {
			PyCapsule_SetName(cap, "used_dltensor");
			if (m->deleter)
				m->deleter(m);
			Py_DECREF(cap);
			return dk_fail(rc, "Unknown error when creating state");
		}
	}
	if (why) {
		Py_DECREF(cap);
		PyErr_SetString(DokiError, why);
		return NULL;
	}
	/* consumed: the values are copied, so the producer's tensor goes now */
	PyCapsule_SetName(cap, "used_dltensor");
	if (m->deleter)
		m->deleter(m);
	Py_DECREF(cap);
	return reg_wrap(st);
}

static PyObject *doki_registry_get(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned long long id;
	int canonical, verbose;
	reg_t *r;
	cplx v;
	(void)self;
	if (!PyArg_ParseTuple(args, "OKpp", &cap, &id, &canonical, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: registry_get(registry, id, canonical, verbose)");
		return NULL;
	}
	r = reg_get(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	if (id >= ((uint64_t)1 << r->nq)) {
		PyErr_SetString(DokiError, "Out of bounds");
		return NULL;
	}
	if (reg_value(r, id, &v))
		return NULL;
	if (canonical) {
		double ph;
		cplx rot;
		if (reg_phase(r, &ph))
			return NULL;
		rot = cos(ph) + I * -sin(ph);
		v = v * rot;
	}
	return PyComplex_FromDoubles(creal(v), cimag(v));
}

/* registry_get_all(registry, canonical, verbose): every amplitude as a 1-D
 * numpy complex128 array whose element i equals registry_get(registry, i,
 * canonical, verbose) bit for bit (same division by the norm constant, same
 * canonical rotation), read with one device-to-host copy.  The reference
 * has no such entry; its callers build the vector with one registry_get
 * call per amplitude (tests/reg_creation_tests.py:24-28). */
static PyObject *doki_registry_get_all(PyObject *self, PyObject *args)
{
	PyObject *cap;
	PyArrayObject *arr;
	int canonical, verbose, rc;
	reg_t *r;
	npy_intp len;
	double *v;
	cplx rot = 1.0;
	(void)self;
	if (!PyArg_ParseTuple(args, "Opp", &cap, &canonical, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: registry_get_all(registry, canonical, verbose)");
		return NULL;
	}
	r = reg_get(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	if (reg_norm(r))
		return NULL;
	if (canonical) {
		double ph;
		if (reg_phase(r, &ph))
			return NULL;
		rot = cos(ph) + I * -sin(ph);
	}
	len = (npy_intp)1 << r->nq;
	arr = (PyArrayObject *)PyArray_SimpleNew(1, &len, NPY_CDOUBLE);
	if (!arr)
		return NULL;
	v = (double *)PyArray_DATA(arr);
	rc = dk_read_raw(r->st, 0, (uint64_t)len, v, NULL);
	if (rc) {
		Py_DECREF(arr);
		dk_fail(rc, "Failed to read state");
		return NULL;
	}
	for (npy_intp i = 0; i < len; i++) {
		cplx x = (v[2 * i] + I * v[2 * i + 1]) / r->norm;
		if (canonical)
			x = x * rot;
		v[2 * i] = creal(x);
		v[2 * i + 1] = cimag(x);
	}
	return (PyObject *)arr;
}

static PyObject *doki_registry_mem(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	reg_t *r;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_mem(state, verbose)");
		return NULL;
	}
	r = reg_peek(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	/* state_mem_size on LP64 (qstate.c:123-135), known before the state is
	 * materialised */
	return PyLong_FromUnsignedLongLong((uint64_t)64 + ((uint64_t)16 << r->nq));
}

/* ------------------------------------------------------------------------ */
/* gates                                                                    */
/* ------------------------------------------------------------------------ */

static void gate_capsule_destroy(PyObject *cap)
{
	gate_t *g = (gate_t *)PyCapsule_GetPointer(cap, CAP_GATE);
	if (!g)
		return;
	if (g->dev)
		dk_gate_free(g->dev);
	PyMem_Free(g->mat);
	PyMem_Free(g);
}

static PyObject *doki_gate_new(PyObject *self, PyObject *args)
{
	PyObject *list, *cap;
	unsigned int nq;
	int verbose;
	gate_t *g;
	Py_ssize_t i, j, dim;
	(void)self;
	if (!PyArg_ParseTuple(args, "IOp", &nq, &list, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: gate_new(num_qubits, gate, verbose)");
		return NULL;
	}
	if (nq == 0) {
		PyErr_SetString(DokiError, "num_qubits can't be zero");
		return NULL;
	}
	if (!PyList_Check(list)) {
		PyErr_SetString(DokiError, "gate must be a list of lists (matrix)");
		return NULL;
	}
	if (nq > DK_MAX_GATE_QUBITS) {
		PyErr_SetString(DokiError, "gates act on at most 6 qubits");
		return NULL;
	}
	dim = (Py_ssize_t)1 << nq;
	if (PyList_Size(list) != dim) {
		PyErr_SetString(DokiError,
				"Wrong matrix size for specified number of qubits");
		return NULL;
	}
	g = (gate_t *)PyMem_Calloc(1, sizeof(gate_t));
	if (!g)
		return PyErr_NoMemory();
	g->k = nq;
	g->dim = (unsigned)dim;
	g->mat = (double *)PyMem_Malloc(2 * dim * dim * sizeof(double));
	if (!g->mat) {
		PyMem_Free(g);
		return PyErr_NoMemory();
	}
	for (i = 0; i < dim; i++) {
		PyObject *row = PyList_GetItem(list, i);
		if (!PyList_Check(row) || PyList_Size(row) != dim) {
			PyErr_SetString(DokiError,
					"rows must be lists of size 2^num_qubits");
			goto fail;
		}
		for (j = 0; j < dim; j++) {
			cplx v;
			if (parse_complex(PyList_GetItem(row, j), &v)) {
				PyErr_SetString(DokiError,
						"matrix elements must be complex numbers");
				goto fail;
			}
			g->mat[2 * (i * dim + j)] = creal(v);
			g->mat[2 * (i * dim + j) + 1] = cimag(v);
		}
	}
	cap = PyCapsule_New(g, CAP_GATE, gate_capsule_destroy);
	if (!cap)
		goto fail;
	return cap;
fail:
	PyMem_Free(g->mat);
	PyMem_Free(g);
	return NULL;
}

static PyObject *doki_gate_get(PyObject *self, PyObject *args)
{
	PyObject *cap, *res;
	int verbose;
	gate_t *g;
	unsigned i, j;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: gate_get(gate, verbose)");
		return NULL;
	}
	g = (gate_t *)PyCapsule_GetPointer(cap, CAP_GATE);
	if (!g) {
		PyErr_SetString(DokiError, "NULL pointer to gate");
		return NULL;
	}
	res = PyList_New(g->dim);
	if (!res)
		return NULL;
	for (i = 0; i < g->dim; i++) {
		PyObject *row = PyList_New(g->dim);
		if (!row) {
			Py_DECREF(res);
			return NULL;
		}
		for (j = 0; j < g->dim; j++)
			PyList_SET_ITEM(row, j,
					PyComplex_FromDoubles(
						g->mat[2 * (i * g->dim + j)],
						g->mat[2 * (i * g->dim + j) + 1]));
		PyList_SET_ITEM(res, i, row);
	}
	return res;
}


/* ------------------------------------------------------------------------ */
/* registry_apply: argument checks in the reference's order and wording     */
/* (doki.c:722-941), then one device launch.                                */
/* ------------------------------------------------------------------------ */

static int pop_qubits(PyObject *set, unsigned count, unsigned *out,
		      const char *typemsg)
{
	PyObject *aux = PySet_New(set);
	unsigned i;
	if (!aux)
		return -1;
	for (i = 0; i < count; i++) {
		PyObject *v = PySet_Pop(aux);
		if (!v || !PyLong_Check(v)) {
			Py_XDECREF(v);
			Py_DECREF(aux);
			PyErr_SetString(DokiError, typemsg);
			return -1;
		}
		out[i] = (unsigned)PyLong_AsLong(v);
		Py_DECREF(v);
	}
	Py_DECREF(aux);
	return 0;
}

static PyObject *doki_registry_apply(PyObject *self, PyObject *args)
{
	PyObject *scap, *gcap, *tlist, *cset, *aset, *aux;
	int nthreads, verbose, rc;
	reg_t *r;
	gate_t *g;
	dk_gate *dg;
	dk_state *out;
	unsigned nt, nc = 0, na = 0, i;
	unsigned targets[DK_MAX_GATE_QUBITS];
	unsigned *ctrl = NULL, *actrl = NULL;
	uint64_t cm = 0, am = 0;
	(void)self;

	if (!PyArg_ParseTuple(args, "OOOOOip", &scap, &gcap, &tlist, &cset, &aset,
			      &nthreads, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: registry_apply(registry, gate, target_list, "
				"control_set, anticontrol_set, num_threads, verbose)");
		return NULL;
	}
	if (check_threads(nthreads))
		return NULL;
	r = reg_peek(scap, "NULL pointer to registry");
	if (!r)
		return NULL;
	g = (gate_t *)PyCapsule_GetPointer(gcap, CAP_GATE);
	if (!g) {
		PyErr_SetString(DokiError, "NULL pointer to gate");
		return NULL;
	}
	if (!PyList_Check(tlist)) {
		PyErr_SetString(DokiError, "target_list must be a list");
		return NULL;
	}
	nt = (unsigned)PyList_Size(tlist);
	if (nt != g->k) {
		PyErr_SetString(DokiError,
				"Wrong number of targets specified for that gate");
		return NULL;
	}
	if (PySet_Check(cset))
		nc = (unsigned)PySet_Size(cset);
	else if (cset != Py_None) {
		PyErr_SetString(DokiError, "control_set must be a set or None");
		return NULL;
	}
	if (PySet_Check(aset))
		na = (unsigned)PySet_Size(aset);
	else if (aset != Py_None) {
		PyErr_SetString(DokiError, "anticontrol_set must be a set or None");
		return NULL;
	}
	if (nc) {
		ctrl = (unsigned *)PyMem_Malloc(nc * sizeof(unsigned));
		if (!ctrl)
			return PyErr_NoMemory();
		if (pop_qubits(cset, nc, ctrl,
			       "control_set must be a set qubit ids (unsigned integers)"))
			goto fail;
		for (i = 0; i < nc; i++) {
			if (ctrl[i] >= r->nq) {
				PyErr_SetString(DokiError, "Control qubit out of range");
				goto fail;
			}
			cm |= (uint64_t)1 << ctrl[i];
		}
	}
	if (na) {
		aux = PySet_New(aset);
		if (!aux)
			goto fail;
		actrl = (unsigned *)PyMem_Malloc(na * sizeof(unsigned));
		if (!actrl) {
			Py_DECREF(aux);
			PyErr_NoMemory();
			goto fail;
		}
		for (i = 0; i < na; i++) {
			PyObject *v = PySet_Pop(aux);
			if (!v || !PyLong_Check(v)) {
				Py_XDECREF(v);
				Py_DECREF(aux);
				PyErr_SetString(DokiError,
						"anticontrol_set must be a set "
						"qubit ids (unsigned integers)");
				goto fail;
			}
			/* PySet_Contains(None, ...) is -1, i.e. true (Q9) */
			if (PySet_Contains(cset, v)) {
				Py_DECREF(v);
				Py_DECREF(aux);
				PyErr_SetString(DokiError,
						"A control cannot also be an anticontrol");
				goto fail;
			}
			actrl[i] = (unsigned)PyLong_AsLong(v);
			Py_DECREF(v);
			if (actrl[i] >= r->nq) {
				Py_DECREF(aux);
				PyErr_SetString(DokiError,
						"Anticontrol qubit out of range");
				goto fail;
			}
			am |= (uint64_t)1 << actrl[i];
		}
		Py_DECREF(aux);
	}
	for (i = 0; i < nt; i++) {
		PyObject *v = PyList_GetItem(tlist, i);
		unsigned j;
		if (!PyLong_Check(v)) {
			PyErr_SetString(DokiError, "target_list must be a list of "
						   "qubit ids (unsigned integers)");
			goto fail;
		}
		if ((nc > 0 && PySet_Contains(cset, v)) ||
		    (na > 0 && PySet_Contains(aset, v))) {
			PyErr_SetString(DokiError, "A target cannot also be a control "
						   "or an anticontrol");
			goto fail;
		}
		targets[i] = (unsigned)PyLong_AsLong(v);
		if (targets[i] >= r->nq) {
			PyErr_SetString(DokiError, "Target qubit out of range");
			goto fail;
		}
		for (j = 0; j < i; j++)
			if (targets[j] == targets[i]) {
				PyErr_SetString(DokiError, "Repeated target qubit");
				goto fail;
			}
	}
	PyMem_Free(ctrl);
	PyMem_Free(actrl);
	ctrl = actrl = NULL;
	if (g_exec == EXEC_EAGER) {
		if (!r->st && reg_materialise(scap))
			return NULL;
		dg = gate_device(g);
		if (!dg)
			return NULL;
		rc = dk_apply(r->st, dg, targets, cm, am, 0, &out);
		if (rc)
			return dk_fail(rc, "Failed to apply gate");
		return reg_wrap(out);
	}
	/* a pending registry: the input's own gate runs first in EXEC_GATES,
	 * and long chains are cut in EXEC_FUSE */
	if (!r->st && (g_exec == EXEC_GATES || r->depth >= FUSE_MAX_PENDING) &&
	    reg_materialise(scap))
		return NULL;
	{
		reg_t *x = (reg_t *)PyMem_Calloc(1, sizeof(reg_t));
		if (!x)
			return PyErr_NoMemory();
		x->nq = r->nq;
		x->parent = scap;
		Py_INCREF(scap);
		x->gate = gcap;
		Py_INCREF(gcap);
		memcpy(x->targets, targets, nt * sizeof(unsigned));
		x->cm = cm;
		x->am = am;
		x->depth = r->st ? 1 : r->depth + 1;
		r->children++;
		return reg_capsule(x);
	}
fail:
	PyMem_Free(ctrl);
	PyMem_Free(actrl);
	return NULL;
}

static PyObject *doki_registry_join(PyObject *self, PyObject *args)
{
	PyObject *c1, *c2;
	int nthreads, verbose, rc;
	reg_t *a, *b;
	dk_state *out;
	(void)self;
	if (!PyArg_ParseTuple(args, "OOip", &c1, &c2, &nthreads, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_join(most_registry, "
					   "least_registry, num_threads, verbose)");
		return NULL;
	}
	if (check_threads(nthreads))
		return NULL;
	a = reg_get(c1, "NULL pointer to registry1");
	if (!a)
		return NULL;
	b = reg_get(c2, "NULL pointer to registry2");
	if (!b)
		return NULL;
	if (a->nq + b->nq > DK_MAX_QUBITS) {
		PyErr_SetString(DokiError, "Number of qubits exceeds maximum");
		return NULL;
	}
	rc = dk_join(a->st, b->st, &out);
	if (rc)
		return dk_fail(rc, "Unknown error when joining states");
	return reg_wrap(out);
}

static PyObject *doki_registry_measure(PyObject *self, PyObject *args)
{
	PyObject *cap, *rolls_obj, *lst, *newcap, *res;
	unsigned long long mask;
	int nthreads, verbose, rc;
	reg_t *r;
	unsigned n, i, m = 0;
	double *rolls = NULL;
	int8_t results[64];
	dk_state *out = NULL;
	(void)self;
	if (!PyArg_ParseTuple(args, "OKOip", &cap, &mask, &rolls_obj, &nthreads,
			      &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_measure(registry, mask, "
					   "roll_list, num_threads, verbose)");
		return NULL;
	}
	if (check_threads(nthreads))
		return NULL;
	r = reg_get(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	if (!PyList_Check(rolls_obj)) {
		PyErr_SetString(DokiError,
				"roll_list must be a list of real numbers in [0, 1)!");
		return NULL;
	}
	n = r->nq;
	for (i = 0; i < n; i++)
		if ((mask >> i) & 1)
			m++;
	rolls = (double *)PyMem_Malloc((m ? m : 1) * sizeof(double));
	if (!rolls)
		return PyErr_NoMemory();
	for (i = 0; i < m; i++) {
		PyObject *o = PyList_GetItem(rolls_obj, i);
		double v = o ? PyFloat_AsDouble(o) : -1.0;
		if (PyErr_Occurred())
			PyErr_Clear();
		if (v < 0 || v >= 1) {  /* NaN passes, as doki.c:1114 */
			PyMem_Free(rolls);
			PyErr_SetString(DokiError, "roll not in interval [0, 1)!");
			return NULL;
		}
		rolls[i] = v;
	}
	rc = dk_measure(r->st, (uint64_t)mask, rolls, m, results, NULL, &out);
	PyMem_Free(rolls);
	if (rc)
		return dk_fail(rc, "Unknown error while collapsing state");
	lst = PyList_New(n);
	if (!lst) {
		if (out)
			dk_state_free(out);
		return NULL;
	}
	for (i = 0; i < n; i++) {
		PyObject *v = results[i] < 0 ? Py_None
					     : (results[i] ? Py_True : Py_False);
		Py_INCREF(v);
		PyList_SET_ITEM(lst, i, v);
	}
	if (out) {
		newcap = reg_wrap(out);
		if (!newcap) {
			Py_DECREF(lst);
			return NULL;
		}
	} else {
		newcap = Py_None;
		Py_INCREF(newcap);
	}
	res = PyTuple_Pack(2, newcap, lst);
	Py_DECREF(newcap);
	Py_DECREF(lst);
	return res;
}

static PyObject *doki_registry_prob(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned int id;
	int nthreads, verbose, rc;
	reg_t *r;
	(void)self;
	if (!PyArg_ParseTuple(args, "OIip", &cap, &id, &nthreads, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_prob(registry, "
					   "qubit_id, num_threads, verbose)");
		return NULL;
	}
	if (check_threads(nthreads))
		return NULL;
	r = reg_get(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	if (id >= r->nq) {
		PyErr_SetString(DokiError, "Qubit id out of range");
		return NULL;
	}
	/* States never change after creation: one pass yields every qubit's
	 * probability, later calls are answered from the cache. */
	if (!r->probs) {
		double *p = (double *)PyMem_Malloc(r->nq * sizeof(double));
		if (!p)
			return PyErr_NoMemory();
		rc = dk_prob_all(r->st, p);
		if (rc) {
			PyMem_Free(p);
			return dk_fail(rc, "Failed to compute probability");
		}
		r->probs = p;
	}
	return PyFloat_FromDouble(r->probs[id]);
}

/* ------------------------------------------------------------------------ */
/* funmatrix: the lazy matrix algebra                                       */
/* ------------------------------------------------------------------------ */
/*
 * A funmatrix is an immutable node: a leaf (custom values, identity, the
 * |0...0> state or density, Walsh/Hadamard, controlled U, I (x) U (x) I, a
 * projection, a partial trace, a gate applied to a state, or the density
 * matrix of a registry) or a binary operation on two funmatrices (sum,
 * difference, matrix product, element-wise product, Kronecker product).
 * Every node carries a scalar factor and transpose / conjugate flags.
 * Element (i, j) is evaluated on demand with the reference's rules
 * (getitem, funmatrix.c:238-340): the bounds are those of the node's own
 * shape, which transposition leaves as it was (funmatrix.c:761-839); then
 * the indices are swapped if the node is transposed, the value computed,
 * conjugated if asked and multiplied by the scalar.  Evaluation codes: 0 ok,
 * 1-5 a parent of a sum / difference / product / element-wise product /
 * Kronecker product failed, 7 out of bounds, 8 NaN, -1 Python error set.
 * Density leaves of registries are evaluated on the GPU.
 * funmatrix_mem reports the bytes the reference's structures would take on
 * LP64 (struct FMatrix 112, Matrix2D 40, Projection 40, DMatrixForTrace 24,
 * Application 72; funmatrix.c:1301-1566, qops.c:295-447).
 */
enum {
	FM_DENSITY = 1, FM_PTRACE, FM_CUSTOM, FM_IDENTITY, FM_STATEZERO,
	FM_DENSITYZERO, FM_WALSH, FM_CU, FM_EYEKRON, FM_PROJECTION, FM_APPLY,
	FM_OP
};
enum { FMOP_ADD = 0, FMOP_SUB, FMOP_MATMUL, FMOP_EWMUL, FMOP_KRON };

#define REF_MATRIX2D_BYTES 40
#define REF_PROJECTION_BYTES 40
#define REF_APPLICATION_BYTES 72

/* Payload shared by a node and its scaled / transposed copies. */
typedef struct {
	Py_ssize_t refs;
	int64_t len;
	cplx *v;              /* custom values at flat index i * rows + j */
	unsigned nt, nc, na;  /* gate application: counts of ... */
	unsigned *ids;        /* ... targets, then controls, then anticontrols */
} fm_buf_t;

typedef struct fm_t {
	int kind;
	int64_t r, c;
	cplx s;               /* scalar factor */
	int transpose, conjugate;
	size_t mem;           /* bytes the reference would report */
	int op;               /* FM_OP */
	PyObject *a_cap;      /* owned: first operand, or the parent / U /
	                         state of a leaf that reads another matrix */
	PyObject *b_cap;      /* owned: second operand, or the applied gate */
	/* FM_DENSITY: reduced density of the kept qubits of a registry */
	PyObject *state_cap;  /* owned reference */
	uint64_t keep_mask;
	unsigned keep_count;
	unsigned char keep_pos[64];
	/* projections folded in: register qubits in pmask must have the bit
	 * of pval -- a row condition for kept qubits, a restriction of the
	 * traced sum for traced ones */
	uint64_t pmask, pval;
	double *cache;        /* materialised matrix, or NULL */
	int e;                /* FM_PTRACE: index bit traced out */
	int hadamard;         /* FM_WALSH */
	int64_t right;        /* FM_EYEKRON: 2^rightQ (the left factor only
	                         shapes the matrix) */
	uint64_t qubit;       /* FM_PROJECTION */
	int value;
	fm_buf_t *buf;        /* FM_CUSTOM, FM_APPLY */
} fm_t;

#define FM_MATERIALISE_MAX_QUBITS 10

static void fm_buf_release(fm_buf_t *b)
{
	if (b && --b->refs == 0) {
		PyMem_Free(b->v);
		PyMem_Free(b->ids);
		PyMem_Free(b);
	}
}

static void fm_free(fm_t *f)
{
	Py_XDECREF(f->a_cap);
	Py_XDECREF(f->b_cap);
	Py_XDECREF(f->state_cap);
	fm_buf_release(f->buf);
	PyMem_Free(f->cache);
	PyMem_Free(f);
}

static void fm_capsule_destroy(PyObject *cap)
{
	fm_t *f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (f)
		fm_free(f);
}

static PyObject *fm_wrap(fm_t *f)
{
	PyObject *cap = PyCapsule_New(f, CAP_FMAT, fm_capsule_destroy);
	if (!cap)
		fm_free(f);
	return cap;
}

static fm_t *fm_new(int kind, int64_t r, int64_t c)
{
	fm_t *f = (fm_t *)PyMem_Calloc(1, sizeof(fm_t));
	if (!f) {
		PyErr_NoMemory();
		return NULL;
	}
	f->kind = kind;
	f->r = r;
	f->c = c;
	f->s = 1.0;
	f->mem = REF_FMATRIX_BYTES;
	return f;
}

/* The node of a funmatrix capsule, or NULL (no error set). */
static fm_t *fm_of(PyObject *cap)
{
	fm_t *f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f)
		PyErr_Clear();
	return f;
}

/* A copy sharing operands and payload (scaled, transposed, conjugated). */
static fm_t *fm_clone(const fm_t *f)
{
	fm_t *o = (fm_t *)PyMem_Malloc(sizeof(fm_t));
	if (!o) {
		PyErr_NoMemory();
		return NULL;
	}
	*o = *f;
	o->cache = NULL;
	Py_XINCREF(o->a_cap);
	Py_XINCREF(o->b_cap);
	Py_XINCREF(o->state_cap);
	if (o->buf)
		o->buf->refs++;
	return o;
}

static PyObject *doki_registry_density(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	reg_t *r;
	fm_t *f;
	unsigned q;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_density(state, verbose)");
		return NULL;
	}
	r = reg_peek(cap, "NULL pointer to registry");
	if (!r)
		return NULL;
	f = fm_new(FM_DENSITY, (int64_t)1 << r->nq, (int64_t)1 << r->nq);
	if (!f)
		return NULL;
	f->mem = REF_FMATRIX_BYTES + (size_t)64 + ((size_t)16 << r->nq);
	f->state_cap = cap;
	Py_INCREF(cap);
	f->keep_count = r->nq;
	for (q = 0; q < r->nq; q++)
		f->keep_pos[q] = (unsigned char)q;
	f->keep_mask = r->nq >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << r->nq) - 1);
	return fm_wrap(f);
}

static int fm_get(fm_t *f, int64_t i, int64_t j, cplx *out);

static int64_t ins_bit(int value, int64_t pos, int bit)
{
	int64_t step = (int64_t)1 << bit;
	return pos % step + (pos / step) * (step << 1) + value * step;
}

/* Whether reduced row index i satisfies the projections on kept qubits. */
static int fm_density_row_ok(const fm_t *f, int64_t i)
{
	unsigned b;
	for (b = 0; b < f->keep_count; b++) {
		const unsigned q = f->keep_pos[b];
		if (((f->pmask >> q) & 1) && ((uint64_t)(i >> b) & 1) != ((f->pval >> q) & 1))
			return 0;
	}
	return 1;
}

static int fm_density_get(fm_t *f, int64_t i, int64_t j, cplx *out)
{
	reg_t *r = reg_get(f->state_cap, "NULL pointer to registry");
	const uint64_t tmask = f->pmask & ~f->keep_mask, tval = f->pval & tmask;
	int rc;
	if (!r)
		return -1;
	if (!fm_density_row_ok(f, i)) {
		*out = 0;
		return 0;
	}
	if (f->keep_count <= FM_MATERIALISE_MAX_QUBITS) {
		if (!f->cache) {
			size_t n = (size_t)f->r * (size_t)f->r;
			f->cache = (double *)PyMem_Malloc(2 * n * sizeof(double));
			if (!f->cache) {
				PyErr_NoMemory();
				return -1;
			}
			rc = dk_reduced_density_masked(r->st, f->keep_mask, tmask, tval,
						       f->cache);
			if (rc) {
				PyMem_Free(f->cache);
				f->cache = NULL;
				dk_fail(rc, "[DENSITY] Unknown error");
				return -1;
			}
		}
		{
			size_t k = (size_t)i * (size_t)f->r + (size_t)j;
			*out = f->cache[2 * k] + I * f->cache[2 * k + 1];
		}
		return 0;
	} else {
		uint64_t ri = (uint64_t)i, cj = (uint64_t)j;
		double v[2];
		rc = dk_reduced_density_elements_masked(r->st, f->keep_mask, tmask, tval,
							1, &ri, &cj, v);
		if (rc) {
			dk_fail(rc, "[DENSITY] Unknown error");
			return -1;
		}
		*out = v[0] + I * v[1];
		return 0;
	}
}

/* Element of another matrix read by a leaf: the code fm_get returns (0,
 * a positive getitem code, or -1 on a Python error) with *out as fm_get
 * left it -- NaN for an index out of bounds, the partial sum of a failed
 * matrix product -- because the reference's leaf functions go on with
 * whatever value their nested getitem produced. */
static int fm_read(PyObject *cap, int64_t i, int64_t j, cplx *out)
{
	fm_t *m = fm_of(cap);
	if (!m) {
		*out = NAN;
		return 8;
	}
	return fm_get(m, i, j, out);
}

/* Materialises an FM_APPLY node: reads its input column and its gate (a
 * failed read stands as NaN, so it spreads to every element that uses it, as
 * in _ApplyGateFunction, qops.c:449-563), applies the gate to the column as
 * one in-place dk_apply on the device, and keeps the raw result (not
 * normalised: the node is a plain matrix-vector product) as f->cache.  A gate
 * on no qubits (1x1) is a scalar on the rows its controls select: it runs as
 * the diagonal gate diag(g, g) on a qubit that is neither control nor
 * anticontrol, or, when there is none, as diag(1, g) / diag(g, 1) on one of
 * the controls / anticontrols; a zero-qubit column is padded to one qubit. */
static int fm_apply_materialise(fm_t *f)
{
	const fm_buf_t *b = f->buf;
	const unsigned *ct = b->ids + b->nt, *ac = b->ids + b->nt + b->nc;
	const int64_t rows = f->r;
	const unsigned nq = rows > 1 ? (unsigned)__builtin_ctzll((uint64_t)rows) : 1;
	const int64_t srows = (int64_t)1 << nq;
	const int64_t dg = (int64_t)1 << b->nt;
	unsigned tg[6], k = b->nt, u;
	uint64_t cm = 0, am = 0;
	double *col = NULL, *mat = NULL, norm;
	dk_state *st = NULL;
	dk_gate *gt = NULL;
	int64_t i, j;
	int rc;
	cplx x;
	if (b->nt > 6) {
		PyErr_SetString(DokiError, "[APPLY] gates act on at most 6 qubits");
		return -1;
	}
	for (u = 0; u < b->nc; u++)
		cm |= (uint64_t)1 << ct[u];
	for (u = 0; u < b->na; u++)
		am |= (uint64_t)1 << ac[u];
	col = (double *)PyMem_Calloc((size_t)srows * 2, sizeof(double));
	mat = (double *)PyMem_Calloc(2 * 64 * 64, sizeof(double));
	if (!col || !mat) {
		PyErr_NoMemory();
		goto fail;
	}
	for (i = 0; i < rows; i++) {
		rc = fm_read(f->a_cap, i, 0, &x);
		if (rc < 0)
			goto fail;
		if (rc)
			x = NAN;
		col[2 * i] = creal(x);
		col[2 * i + 1] = cimag(x);
	}
	for (i = 0; i < dg; i++)
		for (j = 0; j < dg; j++) {
			rc = fm_read(f->b_cap, i, j, &x);
			if (rc < 0)
				goto fail;
			if (rc)
				x = NAN;
			mat[2 * (i * dg + j)] = creal(x);
			mat[2 * (i * dg + j) + 1] = cimag(x);
		}
	for (u = 0; u < b->nt; u++)
		tg[u] = b->ids[u];
	if (b->nt == 0) {
		/* the scalar g = mat[0..1] on a diagonal 2x2 gate */
		const uint64_t used = cm | am;
		double one[2] = {1.0, 0.0}, gv[2] = {mat[0], mat[1]};
		const double *d0 = gv, *d1 = gv;
		k = 1;
		tg[0] = 0;
		while (tg[0] < nq && ((used >> tg[0]) & 1))
			tg[0]++;
		if (tg[0] == nq) {
			/* every qubit is a control or an anticontrol */
			tg[0] = 0;
			if ((cm >> 0) & 1) {
				d0 = one;
				cm &= ~(uint64_t)1;
			} else {
				d1 = one;
				am &= ~(uint64_t)1;
			}
		}
		memset(mat, 0, 8 * sizeof(double));
		mat[0] = d0[0];
		mat[1] = d0[1];
		mat[6] = d1[0];
		mat[7] = d1[1];
	}
	rc = dk_state_from_host(nq, col, &st);
	if (!rc)
		rc = dk_gate_new(k, mat, &gt);
	if (!rc)
		rc = dk_apply(st, gt, tg, cm, am, 1, &st);
	if (!rc)
		rc = dk_read_raw(st, 0, (uint64_t)srows, col, &norm);
	if (rc) {
		dk_fail(rc, "[APPLY] Unknown error");
		goto fail;
	}
	dk_gate_free(gt);
	dk_state_free(st);
	PyMem_Free(mat);
	f->cache = col;
	return 0;
fail:
	if (gt)
		dk_gate_free(gt);
	if (st)
		dk_state_free(st);
	PyMem_Free(col);
	PyMem_Free(mat);
	return -1;
}

/* Value of a leaf at (i, j), indices already swapped for transposition.
 * Follows the reference's element functions: _CustomMat, _IdentityFunction,
 * _StateZeroFunction, _DensityZeroFunction, _WalshFunction, _CUFunction,
 * _eyeKronFunction, _projectionFunction, _PartialTFunct
 * (funmatrix.c:680-1566) and _ApplyGateFunction (qops.c:449-563). */
static int fm_leaf(fm_t *f, int64_t i, int64_t j, cplx *out)
{
	switch (f->kind) {
	case FM_DENSITY:
		return fm_density_get(f, i, j, out);
	case FM_PTRACE: {
		/* _PartialTFunct: the sum of both reads when both succeed; the
		 * first value alone when only the second fails (its index lies
		 * past the matrix); whatever the failed first read left
		 * otherwise */
		cplx a, b;
		int r1 = fm_read(f->a_cap, ins_bit(0, i, f->e), ins_bit(0, j, f->e), &a), r2;
		if (r1 < 0)
			return -1;
		*out = a;
		if (r1)
			return 0;
		r2 = fm_read(f->a_cap, ins_bit(1, i, f->e), ins_bit(1, j, f->e), &b);
		if (r2 < 0)
			return -1;
		if (r2 == 0)
			*out = a + b;
		return 0;
	}
	case FM_CUSTOM: {
		/* flat index with the row count as stride (funmatrix.c:1514);
		 * indices past the stored values read NaN */
		int64_t k = i * f->r + j;
		*out = (k >= 0 && k < f->buf->len) ? f->buf->v[k] : NAN;
		return 0;
	}
	case FM_IDENTITY:
		*out = (double)(i == j) + I * 0.0;
		return 0;
	case FM_STATEZERO:
		*out = (double)(i == 0) + I * 0.0;
		return 0;
	case FM_DENSITYZERO:
		*out = (double)(i == 0 && j == 0) + I * 0.0;
		return 0;
	case FM_WALSH: {
		/* Sylvester order: (-1)^popcount(i & j), over sqrt(size) for
		 * Hadamard */
		double number = (__builtin_popcountll((uint64_t)(i & j)) & 1) ? -1.0 : 1.0;
		if (f->hadamard)
			number /= sqrt((double)f->r);
		*out = number + I * 0.0;
		return 0;
	}
	case FM_CU: {
		fm_t *u = fm_of(f->a_cap);
		if (!u) {
			*out = NAN;
			return 0;
		}
		if (i < u->r || j < u->c) {
			*out = (double)(i == j) + I * 0.0;
			return 0;
		}
		return fm_read(f->a_cap, i - u->r, j - u->c, out) < 0 ? -1 : 0;
	}
	case FM_EYEKRON: {
		fm_t *u = fm_of(f->a_cap);
		if (!u) {
			*out = NAN;
			return 0;
		}
		if (i % f->right != j % f->right) {
			*out = 0;
			return 0;
		}
		i /= f->right;
		j /= f->right;
		if (i / u->r != j / u->c) {
			*out = 0;
			return 0;
		}
		return fm_read(f->a_cap, i % u->r, j % u->c, out) < 0 ? -1 : 0;
	}
	case FM_PROJECTION: {
		const int set = f->qubit < 63 && ((i >> f->qubit) & 1);
		if (set != f->value) {
			*out = 0;
			return 0;
		}
		return fm_read(f->a_cap, i, j, out) < 0 ? -1 : 0;
	}
	case FM_APPLY:
		/* the column is materialised on the device at the first read */
		if (!f->cache && fm_apply_materialise(f) < 0)
			return -1;
		*out = f->cache[2 * i] + I * f->cache[2 * i + 1];
		(void)j;
		return 0;
	default:
		*out = NAN;
		return 0;
	}
}

/* A binary operation node at (i, j). */
static int fm_op(fm_t *f, int64_t i, int64_t j, cplx *out)
{
	fm_t *a = fm_of(f->a_cap), *b = fm_of(f->b_cap);
	static const int fail[] = {1, 2, 3, 4, 5};
	cplx x = NAN, y = NAN;
	int r1, r2;
	int64_t k;
	if (!a || !b)
		return fail[f->op];
	switch (f->op) {
	case FMOP_MATMUL:
		*out = 0;
		for (k = 0; k < a->c; k++) {
			r1 = fm_get(a, i, k, &x);
			r2 = r1 < 0 ? 0 : fm_get(b, k, j, &y);
			if (r1 < 0 || r2 < 0)
				return -1;
			if (r1 || r2)
				return fail[f->op];
			*out = *out + x * y;
		}
		return 0;
	case FMOP_KRON:
		r1 = fm_get(a, i / b->r, j / b->c, &x);
		r2 = r1 < 0 ? 0 : fm_get(b, i % b->r, j % b->c, &y);
		break;
	default:
		r1 = fm_get(a, i, j, &x);
		r2 = r1 < 0 ? 0 : fm_get(b, i, j, &y);
		break;
	}
	if (r1 < 0 || r2 < 0)
		return -1;
	if (r1 || r2)
		return fail[f->op];
	switch (f->op) {
	case FMOP_ADD:
		*out = x + y;
		break;
	case FMOP_SUB:
		*out = x - y;
		break;
	default: /* element-wise and Kronecker products */
		*out = x * y;
		break;
	}
	return 0;
}

static int fm_get(fm_t *f, int64_t i, int64_t j, cplx *out)
{
	int rc;
	*out = NAN;
	if (!(i >= 0 && j >= 0 && i < f->r && j < f->c))
		return 7;
	if (f->transpose) {
		int64_t t = i;
		i = j;
		j = t;
	}
	if (f->kind == FM_OP) {
		rc = fm_op(f, i, j, out);
	} else {
		rc = fm_leaf(f, i, j, out);
		if (rc == 0 && (isnan(creal(*out)) || isnan(cimag(*out))))
			rc = 8;
	}
	if (rc == 0 && f->conjugate)
		*out = conj(*out);
	if (rc == 0)
		*out = *out * f->s;
	return rc;
}

static const char *get_err_text(int code, int trace)
{
	static const char *const get_txt[] = {
		"[GET] Unknown error code",
		"[GET] Error adding parent matrices",
		"[GET] Error substracting parent matrices",
		"[GET] Error multiplying parent matrices",
		"[GET] Error multiplying entity-wise parent matrices",
		"[GET] Error calculating Kronecker product of parent matrices",
		"[GET] Unknown operation between parent matrices",
		"[GET] Element out of bounds",
		"[GET] f returned NAN"};
	static const char *const trace_txt[] = {
		"[TRACE] Unknown error code",
		"[TRACE] Error adding parent matrices",
		"[TRACE] Error substracting parent matrices",
		"[TRACE] Error multiplying parent matrices",
		"[TRACE] Error multiplying entity-wise parent matrices",
		"[TRACE] Error calculating Kronecker product of parent matrices",
		"[TRACE] Unknown operation between parent matrices",
		"[TRACE] Element out of bounds",
		"[TRACE] f returned NAN"};
	if (code < 1 || code > 8)
		code = 0;
	return trace ? trace_txt[code] : get_txt[code];
}

static PyObject *doki_funmatrix_get(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned long long i, j;
	int verbose, rc;
	fm_t *f;
	cplx v;
	(void)self;
	if (!PyArg_ParseTuple(args, "OKKp", &cap, &i, &j, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: funmatrix_get(funmatrix, i, j, verbose)");
		return NULL;
	}
	f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f) {
		PyErr_SetString(DokiError, "NULL pointer to matrix");
		return NULL;
	}
	if ((int64_t)i < 0 || (int64_t)j < 0 || (int64_t)i >= f->r ||
	    (int64_t)j >= f->c) {
		PyErr_SetString(DokiError, "Out of bounds");
		return NULL;
	}
	rc = fm_get(f, (int64_t)i, (int64_t)j, &v);
	if (rc < 0)
		return NULL;
	if (rc) {
		PyErr_SetString(DokiError, get_err_text(rc, 0));
		return NULL;
	}
	if (isnan(creal(v)) || isnan(cimag(v))) {
		PyErr_SetString(DokiError, "[GET] Unexpected NAN value");
		return NULL;
	}
	return PyComplex_FromDoubles(creal(v), cimag(v));
}

static PyObject *doki_funmatrix_shape(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	fm_t *f;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_shape(funmatrix, verbose)");
		return NULL;
	}
	f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f) {
		PyErr_SetString(DokiError, "NULL pointer to matrix");
		return NULL;
	}
	return Py_BuildValue("(KK)", (unsigned long long)f->r,
			     (unsigned long long)f->c);
}

static PyObject *doki_funmatrix_partialtrace(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned int id;
	int verbose, e;
	fm_t *f, *o;
	(void)self;
	if (!PyArg_ParseTuple(args, "OIp", &cap, &id, &verbose)) {
		PyErr_SetString(DokiError,
				"Syntax: funmatrix_partialtrace(funmatrix, id, verbose)");
		return NULL;
	}
	f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f) {
		PyErr_SetString(DokiError, "[PTRACE] The matrix is NULL");
		return NULL;
	}
	if (f->r != f->c) {
		PyErr_SetString(DokiError, "[PTRACE] The matrix is not a square matrix");
		return NULL;
	}
	e = (int)id;
	if (e < 0) {
		PyErr_SetString(DokiError, "[PTRACE] elem id has to be >= 0");
		return NULL;
	}
	o = fm_new(FM_PTRACE, f->r >> 1, f->c >> 1);
	if (!o)
		return NULL;
	o->mem = REF_FMATRIX_BYTES + REF_PTRACE_ARGV_BYTES + f->mem;
	if (f->kind == FM_DENSITY && (unsigned)e < f->keep_count &&
	    f->s == 1.0 && !f->transpose && !f->conjugate) {
		/* trace out the e-th remaining qubit of a registry density:
		 * the same node kind with one fewer kept qubit (higher ones
		 * shift down, funmatrix.c:905-918), evaluated on the GPU */
		unsigned k, w = 0;
		o->kind = FM_DENSITY;
		o->state_cap = f->state_cap;
		Py_INCREF(o->state_cap);
		for (k = 0; k < f->keep_count; k++)
			if ((int)k != e)
				o->keep_pos[w++] = f->keep_pos[k];
		o->keep_count = w;
		o->keep_mask = f->keep_mask & ~((uint64_t)1 << f->keep_pos[e]);
		o->pmask = f->pmask;
		o->pval = f->pval;
	} else {
		o->a_cap = cap;
		Py_INCREF(cap);
		o->e = e;
	}
	return fm_wrap(o);
}

static PyObject *doki_funmatrix_trace(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	fm_t *f;
	int64_t i, n;
	cplx acc = 0, v;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_trace(funmatrix, verbose)");
		return NULL;
	}
	f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f) {
		PyErr_SetString(DokiError, "NULL pointer to matrix");
		return NULL;
	}
	n = f->r <= f->c ? f->r : f->c;
	for (i = 0; i < n; i++) {
		int rc = fm_get(f, i, i, &v);
		if (rc < 0)
			return NULL;
		if (rc) {
			PyErr_SetString(DokiError, get_err_text(rc, 1));
			return NULL;
		}
		if (isnan(creal(v)) || isnan(cimag(v))) {
			PyErr_SetString(DokiError, "[TRACE] Unexpected NAN value");
			return NULL;
		}
		acc = acc + v;
	}
	return PyComplex_FromDoubles(creal(acc), cimag(acc));
}

static PyObject *doki_funmatrix_mem(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	fm_t *f;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_mem(fmatrix, verbose)");
		return NULL;
	}
	f = (fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT);
	if (!f) {
		PyErr_SetString(DokiError, "NULL pointer to FMatrix");
		return NULL;
	}
	return PyLong_FromSize_t(f->mem);
}

/* ---- constructors: doki.c:1297-1497 ------------------------------------ */

/* A number argument (complex, float or int) as the reference reads one. */
static int fm_number(PyObject *o, cplx *out)
{
	if (PyComplex_Check(o))
		*out = PyComplex_RealAsDouble(o) + I * PyComplex_ImagAsDouble(o);
	else if (PyFloat_Check(o))
		*out = PyFloat_AsDouble(o) + I * 0.0;
	else if (PyLong_Check(o))
		*out = (double)PyLong_AsLong(o) + I * 0.0;
	else
		return -1;
	return 0;
}

static PyObject *doki_funmatrix_create(PyObject *self, PyObject *args)
{
	PyObject *list, *row;
	Py_ssize_t nr, nc, i, j;
	int verbose;
	fm_buf_t *b;
	fm_t *f;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &list, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_create(matrix, verbose)");
		return NULL;
	}
	if (!PyList_Check(list)) {
		PyErr_SetString(DokiError, "matrix must be a list of lists (matrix)");
		return NULL;
	}
	nr = PyList_Size(list);
	if (nr == 0) {
		PyErr_SetString(DokiError, "there are no rows");
		return NULL;
	}
	row = PyList_GetItem(list, 0);
	if (!PyList_Check(row)) {
		PyErr_SetString(DokiError, "rows must be lists");
		return NULL;
	}
	nc = PyList_Size(row);
	if (nc == 0) {
		PyErr_SetString(DokiError, "there are no columns");
		return NULL;
	}
	b = (fm_buf_t *)PyMem_Calloc(1, sizeof(fm_buf_t));
	if (!b)
		return PyErr_NoMemory();
	b->refs = 1;
	b->len = (int64_t)nr * nc;
	b->v = (cplx *)PyMem_Calloc((size_t)b->len, sizeof(cplx));
	if (!b->v) {
		fm_buf_release(b);
		return PyErr_NoMemory();
	}
	for (i = 0; i < nr; i++) {
		row = PyList_GetItem(list, i);
		if (!PyList_Check(row) || PyList_Size(row) != nc) {
			PyErr_SetString(DokiError, "rows must be lists of the same size");
			fm_buf_release(b);
			return NULL;
		}
		for (j = 0; j < nc; j++) {
			cplx v;
			int64_t k = (int64_t)i * nr + j;
			if (fm_number(PyList_GetItem(row, j), &v)) {
				PyErr_SetString(DokiError,
						"matrix elements must be complex numbers");
				fm_buf_release(b);
				return NULL;
			}
			/* the reference's layout (row stride = row count,
			 * doki.c:1368); values that would land past the array
			 * are dropped (the reference writes them out of bounds) */
			if (k < b->len)
				b->v[k] = v;
		}
	}
	f = fm_new(FM_CUSTOM, nr, nc);
	if (!f) {
		fm_buf_release(b);
		return NULL;
	}
	f->buf = b;
	f->mem = REF_FMATRIX_BYTES + REF_MATRIX2D_BYTES + (size_t)b->len * sizeof(cplx);
	return fm_wrap(f);
}

/* identity, statezero, densityzero, hadamard: 2^n-sized leaves */
static PyObject *fm_sized_leaf(PyObject *args, const char *syntax, int kind)
{
	unsigned int n;
	int verbose;
	fm_t *f;
	int64_t size;
	if (!PyArg_ParseTuple(args, "Ip", &n, &verbose)) {
		PyErr_SetString(DokiError, syntax);
		return NULL;
	}
	if (n > 62) {
		PyErr_SetString(DokiError, "Number of qubits exceeds maximum");
		return NULL;
	}
	size = (int64_t)1 << n;
	f = fm_new(kind, size, kind == FM_STATEZERO ? 1 : size);
	if (!f)
		return NULL;
	return fm_wrap(f);
}

static PyObject *doki_funmatrix_identity(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_sized_leaf(args, "Syntax: funmatrix_identity(num_qubits, verbose)",
			     FM_IDENTITY);
}

static PyObject *doki_funmatrix_statezero(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_sized_leaf(args, "Syntax: funmatrix_statezero(num_qubits, verbose)",
			     FM_STATEZERO);
}

static PyObject *doki_funmatrix_densityzero(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_sized_leaf(args,
			     "Syntax: funmatrix_densityzero(num_qubits, verbose)",
			     FM_DENSITYZERO);
}

static PyObject *doki_funmatrix_hadamard(PyObject *self, PyObject *args)
{
	PyObject *cap;
	(void)self;
	cap = fm_sized_leaf(args, "Syntax: funmatrix_hadamard(num_qubits, verbose)",
			    FM_WALSH);
	if (cap)
		((fm_t *)PyCapsule_GetPointer(cap, CAP_FMAT))->hadamard = 1;
	return cap;
}

static PyObject *doki_funmatrix_addcontrol(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose;
	fm_t *u, *f;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_addcontrol(funmatrix, verbose)");
		return NULL;
	}
	u = fm_of(cap);
	if (!u) {
		PyErr_SetString(DokiError, "NULL pointer to matrix");
		return NULL;
	}
	f = fm_new(FM_CU, u->r * 2, u->c * 2);
	if (!f)
		return NULL;
	f->a_cap = cap;
	Py_INCREF(cap);
	f->mem = REF_FMATRIX_BYTES + u->mem;
	return fm_wrap(f);
}

/* ---- operations: doki.c:1581-2055 ---------------------------------------- */

/* Binary operation node; NULL with doki.error set as the reference's
 * wrappers do (tag: "ADD", "SUB", "MATMUL", "EWMUL", "KRON"). */
static PyObject *fm_binary(PyObject *ca, PyObject *cb, int op, const char *tag)
{
	fm_t *a = fm_of(ca), *b = fm_of(cb), *f;
	int64_t r, c;
	char msg[96];
	if (!a || !b) {
		snprintf(msg, sizeof(msg), "[%s] The %s operand is NULL", tag,
			 !a ? "first" : "second");
		PyErr_SetString(DokiError, msg);
		return NULL;
	}
	switch (op) {
	case FMOP_MATMUL:
		if (a->c != b->r)
			goto misaligned;
		r = a->r;
		c = b->c;
		break;
	case FMOP_KRON:
		r = a->r * b->r;
		c = a->c * b->c;
		break;
	case FMOP_EWMUL:
		if (a->r == b->r && a->c == b->c) {
			r = a->r;
			c = a->c;
			break;
		}
		/* row .* column and column .* row are outer products
		 * (funmatrix.c:580-584) */
		if (a->r == 1 && b->c == 1)
			return fm_binary(cb, ca, FMOP_MATMUL, tag);
		if (b->r == 1 && a->c == 1)
			return fm_binary(ca, cb, FMOP_MATMUL, tag);
		goto misaligned;
	default:
		if (a->r != b->r || a->c != b->c)
			goto misaligned;
		r = a->r;
		c = a->c;
		break;
	}
	f = fm_new(FM_OP, r, c);
	if (!f)
		return NULL;
	f->op = op;
	f->a_cap = ca;
	f->b_cap = cb;
	Py_INCREF(ca);
	Py_INCREF(cb);
	f->mem = REF_FMATRIX_BYTES + a->mem + b->mem;
	return fm_wrap(f);
misaligned:
	snprintf(msg, sizeof(msg), "[%s] The operands are misalligned", tag);
	PyErr_SetString(DokiError, msg);
	return NULL;
}

#define FM_BINARY_ENTRY(name, op, tag)                                        \
	static PyObject *doki_funmatrix_##name(PyObject *self, PyObject *args)  \
	{                                                                     \
		PyObject *a, *b;                                              \
		int verbose;                                                  \
		(void)self;                                                   \
		if (!PyArg_ParseTuple(args, "OOp", &a, &b, &verbose)) {       \
			PyErr_SetString(DokiError, "Syntax: funmatrix_" #name \
					"(funmatrix1, funmatrix2, verbose)"); \
			return NULL;                                          \
		}                                                             \
		return fm_binary(a, b, op, tag);                              \
	}

FM_BINARY_ENTRY(add, FMOP_ADD, "ADD")
FM_BINARY_ENTRY(sub, FMOP_SUB, "SUB")
FM_BINARY_ENTRY(matmul, FMOP_MATMUL, "MATMUL")
FM_BINARY_ENTRY(ewmul, FMOP_EWMUL, "EWMUL")
FM_BINARY_ENTRY(kron, FMOP_KRON, "KRON")

/* scalar_mul (div = 0) and scalar_div (div = 1): a scaled copy */
static PyObject *fm_scaled(PyObject *args, int div)
{
	PyObject *cap, *raw;
	int verbose;
	cplx x;
	fm_t *m, *f;
	if (!PyArg_ParseTuple(args, "OOp", &cap, &raw, &verbose)) {
		PyErr_SetString(DokiError, div ? "Syntax: funmatrix_div(funmatrix, scalar, verbose)"
					       : "Syntax: funmatrix_mul(funmatrix, scalar, verbose)");
		return NULL;
	}
	if (fm_number(raw, &x)) {
		PyErr_SetString(DokiError, "scalar is not a number");
		return NULL;
	}
	if (div && creal(x) == 0 && cimag(x) == 0) {
		PyErr_SetString(DokiError, "Dividing by zero");
		return NULL;
	}
	m = fm_of(cap);
	if (!m) {
		PyErr_SetString(DokiError, div ? "[SDIV] The matrix operand is NULL"
					       : "[SPROD] The matrix operand is NULL");
		return NULL;
	}
	f = fm_clone(m);
	if (!f)
		return NULL;
	f->s = div ? m->s / x : m->s * x;
	return fm_wrap(f);
}

static PyObject *doki_funmatrix_scalar_mul(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_scaled(args, 0);
}

static PyObject *doki_funmatrix_scalar_div(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_scaled(args, 1);
}

static PyObject *doki_funmatrix_eyekron(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned int left, right;
	int verbose;
	fm_t *u, *f;
	int64_t l, rr;
	(void)self;
	if (!PyArg_ParseTuple(args, "OIIp", &cap, &left, &right, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_eyekron(funmatrix, "
					   "leftQubits, rightQubits, verbose)");
		return NULL;
	}
	u = fm_of(cap);
	if (!u) {
		PyErr_SetString(DokiError, "[EYEKRON] The matrix is NULL");
		return NULL;
	}
	if (left > 62 || right > 62) {
		PyErr_SetString(DokiError, "Number of qubits exceeds maximum");
		return NULL;
	}
	l = (int64_t)1 << left;
	rr = (int64_t)1 << right;
	f = fm_new(FM_EYEKRON, l * u->r * rr, l * u->c * rr);
	if (!f)
		return NULL;
	f->a_cap = cap;
	Py_INCREF(cap);
	f->right = rr;
	f->mem = REF_FMATRIX_BYTES + REF_MATRIX2D_BYTES + u->mem + 2 * sizeof(int64_t);
	return fm_wrap(f);
}

/* transpose (conj = 0) and dagger (conj = 1): a flipped copy, same shape */
static PyObject *fm_flipped(PyObject *args, int conj_too)
{
	PyObject *cap;
	int verbose;
	fm_t *m, *f;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, conj_too ? "Syntax: funmatrix_dagger(funmatrix, verbose)"
						    : "Syntax: funmatrix_transpose(funmatrix, verbose)");
		return NULL;
	}
	m = fm_of(cap);
	if (!m) {
		PyErr_SetString(DokiError, conj_too ? "[HTRANS] The matrix is NULL"
						    : "[TRANS] The matrix is NULL");
		return NULL;
	}
	f = fm_clone(m);
	if (!f)
		return NULL;
	f->transpose = !m->transpose;
	if (conj_too)
		f->conjugate = !m->conjugate;
	return fm_wrap(f);
}

static PyObject *doki_funmatrix_transpose(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_flipped(args, 0);
}

static PyObject *doki_funmatrix_dagger(PyObject *self, PyObject *args)
{
	(void)self;
	return fm_flipped(args, 1);
}

static PyObject *doki_funmatrix_projection(PyObject *self, PyObject *args)
{
	PyObject *cap;
	unsigned int qubit;
	int value, verbose;
	fm_t *m, *f;
	(void)self;
	if (!PyArg_ParseTuple(args, "OIpp", &cap, &qubit, &value, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_projection(funmatrix, "
					   "qubit_id, value, verbose)");
		return NULL;
	}
	m = fm_of(cap);
	if (!m) {
		PyErr_SetString(DokiError, "[PROJ] The matrix is NULL");
		return NULL;
	}
	f = fm_new(FM_PROJECTION, m->r, m->c);
	if (!f)
		return NULL;
	f->mem = REF_FMATRIX_BYTES + REF_PROJECTION_BYTES + m->mem;
	if (m->kind == FM_DENSITY && qubit < m->keep_count && m->s == 1.0 &&
	    !m->transpose && !m->conjugate &&
	    !(((m->pmask >> m->keep_pos[qubit]) & 1) &&
	      ((m->pval >> m->keep_pos[qubit]) & 1) != (uint64_t)(value != 0))) {
		/* a projection of a registry density: the same leaf with one
		 * more bit condition, evaluated on the GPU (a condition that
		 * contradicts an earlier one stays a node of its own) */
		const unsigned q = m->keep_pos[qubit];
		unsigned k;
		f->kind = FM_DENSITY;
		f->state_cap = m->state_cap;
		Py_INCREF(f->state_cap);
		f->keep_count = m->keep_count;
		for (k = 0; k < m->keep_count; k++)
			f->keep_pos[k] = m->keep_pos[k];
		f->keep_mask = m->keep_mask;
		f->pmask = m->pmask | ((uint64_t)1 << q);
		f->pval = (m->pval & ~((uint64_t)1 << q)) |
			  ((uint64_t)(value != 0) << q);
		return fm_wrap(f);
	}
	f->a_cap = cap;
	Py_INCREF(cap);
	f->qubit = qubit;
	f->value = value != 0;
	return fm_wrap(f);
}

/* floor(log2(x)) for x >= 1 */
static unsigned fm_log2(int64_t x)
{
	unsigned k = 0;
	while (x > 1) {
		x >>= 1;
		k++;
	}
	return k;
}

static PyObject *doki_funmatrix_apply(PyObject *self, PyObject *args)
{
	PyObject *scap, *gcap, *tlist, *cset, *aset;
	int verbose;
	fm_t *st, *g, *f;
	unsigned nq, nqg, nt, nc = 0, na = 0, i;
	fm_buf_t *b;
	(void)self;
	if (!PyArg_ParseTuple(args, "OOOOOp", &scap, &gcap, &tlist, &cset, &aset,
			      &verbose)) {
		PyErr_SetString(DokiError, "Syntax: funmatrix_apply(registry, gate, "
					   "target_list, control_set, anticontrol_set, "
					   "verbose)");
		return NULL;
	}
	st = fm_of(scap);
	if (!st) {
		PyErr_SetString(DokiError, "NULL pointer to registry");
		return NULL;
	}
	g = fm_of(gcap);
	if (!g) {
		PyErr_SetString(DokiError, "NULL pointer to gate");
		return NULL;
	}
	if (st->c > 1) {
		PyErr_SetString(DokiError, "registry is not a column vector");
		return NULL;
	}
	if (g->c != g->r) {
		PyErr_SetString(DokiError, "gates have to be square matrices");
		return NULL;
	}
	if (g->r > st->r) {
		PyErr_SetString(DokiError, "gate is too big for this state vector");
		return NULL;
	}
	nq = fm_log2(st->r);
	if (st->r != (int64_t)1 << nq) {
		PyErr_SetString(DokiError, "registry needs 2^n rows");
		return NULL;
	}
	nqg = fm_log2(g->r);
	if (g->r != (int64_t)1 << nqg) {
		PyErr_SetString(DokiError, "gates need 2^n x 2^n elements");
		return NULL;
	}
	if (!PyList_Check(tlist)) {
		PyErr_SetString(DokiError, "target_list must be a list");
		return NULL;
	}
	nt = (unsigned)PyList_Size(tlist);
	if (nt != nqg) {
		PyErr_SetString(DokiError, "Wrong number of targets specified for that gate");
		return NULL;
	}
	if (PySet_Check(cset))
		nc = (unsigned)PySet_Size(cset);
	else if (cset != Py_None) {
		PyErr_SetString(DokiError, "control_set must be a set or None");
		return NULL;
	}
	if (PySet_Check(aset))
		na = (unsigned)PySet_Size(aset);
	else if (aset != Py_None) {
		PyErr_SetString(DokiError, "anticontrol_set must be a set or None");
		return NULL;
	}
	b = (fm_buf_t *)PyMem_Calloc(1, sizeof(fm_buf_t));
	if (!b)
		return PyErr_NoMemory();
	b->refs = 1;
	b->nt = nt;
	b->nc = nc;
	b->na = na;
	b->ids = (unsigned *)PyMem_Calloc((size_t)nt + nc + na + 1, sizeof(unsigned));
	if (!b->ids) {
		fm_buf_release(b);
		return PyErr_NoMemory();
	}
	if (nc) {
		unsigned *ct = b->ids + nt;
		if (pop_qubits(cset, nc, ct,
			       "control_set must be a set qubit ids (unsigned integers)"))
			goto fail;
		for (i = 0; i < nc; i++)
			if (ct[i] >= nq) {
				PyErr_SetString(DokiError, "Control qubit out of range");
				goto fail;
			}
	}
	if (na) {
		/* in the reference's order: type, then "control and
		 * anticontrol" (a None control set counts as containing
		 * everything, doki.c:2370), then range */
		PyObject *aux = PySet_New(aset);
		unsigned *ac = b->ids + nt + nc;
		if (!aux)
			goto fail;
		for (i = 0; i < na; i++) {
			PyObject *v = PySet_Pop(aux);
			if (!v || !PyLong_Check(v)) {
				Py_XDECREF(v);
				Py_DECREF(aux);
				PyErr_SetString(DokiError, "anticontrol_set must be a set "
							   "qubit ids (unsigned integers)");
				goto fail;
			}
			if (cset == Py_None || PySet_Contains(cset, v)) {
				Py_DECREF(v);
				Py_DECREF(aux);
				PyErr_SetString(DokiError,
						"A control cannot also be an anticontrol");
				goto fail;
			}
			ac[i] = (unsigned)PyLong_AsLong(v);
			Py_DECREF(v);
			if (ac[i] >= nq) {
				Py_DECREF(aux);
				PyErr_SetString(DokiError, "Anticontrol qubit out of range");
				goto fail;
			}
		}
		Py_DECREF(aux);
	}
	for (i = 0; i < nt; i++) {
		PyObject *v = PyList_GetItem(tlist, i);
		if (!PyLong_Check(v)) {
			PyErr_SetString(DokiError, "target_list must be a list of qubit "
						   "ids (unsigned integers)");
			goto fail;
		}
		if ((nc > 0 && PySet_Contains(cset, v)) ||
		    (na > 0 && PySet_Contains(aset, v))) {
			PyErr_SetString(DokiError,
					"A target cannot also be a control or an anticontrol");
			goto fail;
		}
		b->ids[i] = (unsigned)PyLong_AsLong(v);
		if (b->ids[i] >= nq) {
			PyErr_SetString(DokiError, "Target qubit out of range");
			goto fail;
		}
	}
	f = fm_new(FM_APPLY, st->r, 1);
	if (!f)
		goto fail;
	f->a_cap = scap;
	f->b_cap = gcap;
	Py_INCREF(scap);
	Py_INCREF(gcap);
	f->buf = b;
	f->mem = REF_FMATRIX_BYTES + REF_APPLICATION_BYTES + st->mem + g->mem +
		 (size_t)(nt + nc + na) * sizeof(unsigned);
	return fm_wrap(f);
fail:
	fm_buf_release(b);
	return NULL;
}


/* ------------------------------------------------------------------------ */

/* execution_mode(mode, verbose): how registry_apply runs its gates (see
 * EXEC_FUSE): "fuse" (chains run as cache-blocked passes when read, the
 * default), "gates" (one gate per call, in place when its input is no
 * longer referenced) or "eager" (each call launches its gate at once, out of
 * place, as the reference does).  None only queries.  Returns the mode in
 * force before the call.  Registries made under one mode read the same
 * values under any other.  No reference counterpart. */
static PyObject *doki_execution_mode(PyObject *self, PyObject *args)
{
	static const char *names[] = {"eager", "gates", "fuse"};
	PyObject *m;
	int verbose = 0, prev = g_exec;
	(void)self;
	if (!PyArg_ParseTuple(args, "O|p", &m, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: execution_mode(mode, verbose)");
		return NULL;
	}
	if (m != Py_None) {
		const char *v = PyUnicode_Check(m) ? PyUnicode_AsUTF8(m) : NULL;
		int k;
		for (k = 0; v && k < 3; k++)
			if (!strcmp(v, names[k]))
				break;
		if (!v || k == 3) {
			PyErr_Clear();
			PyErr_SetString(DokiError,
					"mode must be \"fuse\", \"gates\", \"eager\" or None");
			return NULL;
		}
		g_exec = k;
	}
	return PyUnicode_FromString(names[prev]);
}

/* registry_sync(registry, verbose): runs every pending gate of the registry
 * and waits until the device has finished the work queued so far (for
 * timing; reads do not need it).  No reference counterpart: the reference
 * computes in the call. */
static PyObject *doki_registry_sync(PyObject *self, PyObject *args)
{
	PyObject *cap;
	int verbose, rc;
	(void)self;
	if (!PyArg_ParseTuple(args, "Op", &cap, &verbose)) {
		PyErr_SetString(DokiError, "Syntax: registry_sync(registry, verbose)");
		return NULL;
	}
	if (!reg_get(cap, "NULL pointer to registry"))
		return NULL;
	rc = dk_synchronize();
	if (rc)
		return dk_fail(rc, "Failed to synchronise");
	Py_RETURN_NONE;
}

static PyMethodDef DokiMethods[] = {
	{"gate_new", doki_gate_new, METH_VARARGS, "Create new gate"},
	{"gate_get", doki_gate_get, METH_VARARGS, "Get matrix associated to gate"},
	{"registry_new", doki_registry_new, METH_VARARGS, "Create new registry"},
	{"registry_new_data", doki_registry_new_data, METH_VARARGS,
	 "Create new registry initialized with the specified values"},
	{"registry_clone", doki_registry_clone, METH_VARARGS, "Clone a registry"},
	{"registry_del", doki_registry_del, METH_VARARGS, "Destroy a registry"},
	{"registry_get", doki_registry_get, METH_VARARGS, "Get value from registry"},
	{"registry_get_all", doki_registry_get_all, METH_VARARGS,
	 "Get every value of a registry as a numpy array (one copy)"},
	{"registry_to_dlpack", doki_registry_to_dlpack, METH_VARARGS,
	 "The values of a registry as a DLPack tensor on the device (a copy)"},
	{"registry_from_dlpack", doki_registry_from_dlpack, METH_VARARGS,
	 "A registry holding the values of a DLPack tensor"},
	{"registry_apply", doki_registry_apply, METH_VARARGS, "Apply a gate"},
	{"registry_join", doki_registry_join, METH_VARARGS, "Merges two registries"},
	{"registry_measure", doki_registry_measure, METH_VARARGS,
	 "Measures and collapses specified qubits"},
	{"registry_prob", doki_registry_prob, METH_VARARGS,
	 "Get the chances of obtaining 1 when measuring a certain qubit"},
	{"registry_density", doki_registry_density, METH_VARARGS,
	 "Get the density matrix"},
	{"registry_mem", doki_registry_mem, METH_VARARGS,
	 "Get the memory allocated by this registry in bytes"},
	{"registry_sync", doki_registry_sync, METH_VARARGS,
	 "Run the registry's pending gates and wait for the device"},
	{"execution_mode", doki_execution_mode, METH_VARARGS,
	 "Get or set how registry_apply runs gates: fuse, gates or eager"},
	{"funmatrix_create", doki_funmatrix_create, METH_VARARGS,
	 "Create a functional matrix from a matrix"},
	{"funmatrix_identity", doki_funmatrix_identity, METH_VARARGS,
	 "Create an identity functional matrix"},
	{"funmatrix_statezero", doki_funmatrix_statezero, METH_VARARGS,
	 "Create a functional matrix of the state vector |0>"},
	{"funmatrix_densityzero", doki_funmatrix_densityzero, METH_VARARGS,
	 "Create a functional matrix of the density matrix |0><0|"},
	{"funmatrix_hadamard", doki_funmatrix_hadamard, METH_VARARGS,
	 "Create a functional matrix of a Hadamard gate"},
	{"funmatrix_addcontrol", doki_funmatrix_addcontrol, METH_VARARGS,
	 "Add a control qubit to a gate"},
	{"funmatrix_get", doki_funmatrix_get, METH_VARARGS,
	 "Get a value from a functional matrix"},
	{"funmatrix_add", doki_funmatrix_add, METH_VARARGS, "Matrix addition"},
	{"funmatrix_sub", doki_funmatrix_sub, METH_VARARGS, "Matrix subtraction"},
	{"funmatrix_scalar_mul", doki_funmatrix_scalar_mul, METH_VARARGS,
	 "Scalar product"},
	{"funmatrix_scalar_div", doki_funmatrix_scalar_div, METH_VARARGS,
	 "Scalar division"},
	{"funmatrix_matmul", doki_funmatrix_matmul, METH_VARARGS,
	 "Matrix product"},
	{"funmatrix_ewmul", doki_funmatrix_ewmul, METH_VARARGS,
	 "Entity-wise product"},
	{"funmatrix_kron", doki_funmatrix_kron, METH_VARARGS, "Kronecker product"},
	{"funmatrix_eyekron", doki_funmatrix_eyekron, METH_VARARGS,
	 "Kronecker product of I(2^left), U and I(2^right)"},
	{"funmatrix_transpose", doki_funmatrix_transpose, METH_VARARGS,
	 "Transpose"},
	{"funmatrix_dagger", doki_funmatrix_dagger, METH_VARARGS,
	 "Conjugate transpose"},
	{"funmatrix_projection", doki_funmatrix_projection, METH_VARARGS,
	 "Projection over a column vector"},
	{"funmatrix_shape", doki_funmatrix_shape, METH_VARARGS,
	 "Get a tuple with the shape of the matrix"},
	{"funmatrix_partialtrace", doki_funmatrix_partialtrace, METH_VARARGS,
	 "Get the partial trace of a functional matrix"},
	{"funmatrix_trace", doki_funmatrix_trace, METH_VARARGS,
	 "Get the trace of a functional matrix"},
	{"funmatrix_apply", doki_funmatrix_apply, METH_VARARGS,
	 "Apply a gate to a functional state vector"},
	{"funmatrix_mem", doki_funmatrix_mem, METH_VARARGS,
	 "Get the memory allocated by this FMatrix in bytes"},
	{NULL, NULL, 0, NULL}};

static struct PyModuleDef dokimodule = {PyModuleDef_HEAD_INIT, "doki",
					"MI355X state-vector core (drop-in doki)",
					-1, DokiMethods};

static void doki_atexit(void)
{
	dk_shutdown();
}

PyMODINIT_FUNC PyInit_doki(void)
{
	PyObject *m;
	import_array();
	m = PyModule_Create(&dokimodule);
	if (!m)
		return NULL;
	DokiError = PyErr_NewException("qsimov.doki.error", NULL, NULL);
	Py_XINCREF(DokiError);
	if (PyModule_AddObject(m, "error", DokiError) < 0) {
		Py_XDECREF(DokiError);
		Py_CLEAR(DokiError);
		Py_DECREF(m);
		return NULL;
	}
	Py_AtExit(doki_atexit);
	return m;
}
