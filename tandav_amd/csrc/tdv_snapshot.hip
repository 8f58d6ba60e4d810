/*
 * Gadget format-2 snapshots (IO/write_snapshot.c, IO/read_snapshot.c,
 * block table IO/io.h:68-89) for one rank, on the host over the particle
 * columns the context holds.
 *
 * A file is a sequence of blocks, each two Fortran records: the 8-byte
 * label record (4 label characters, the data record's size + 8) and the
 * data record (write_block_header, write_snapshot.c:276-297).  A Fortran
 * record is its payload between two 4-byte copies of its length.  Per
 * particle block the values of each particle are stored together (x, y, z
 * of one particle, then the next; fill_data_buffer, write_snapshot.c:
 * 221-270).  With DOUBLE_PRECISION and without LONG_IDS (the engine's
 * configuration): POS, VEL 3 x double, ID uint32, MASS double, ACC one
 * double (the block table declares one component of P.Acc, so it holds
 * the x component, io.h:76), GACC 3 x double.
 */
#include "tdv_internal.h"

#include <cstdio>
#include <cstring>
#include <vector>

static_assert(sizeof(tdv_snapshot_header) == 256,
		"the Gadget header is 256 bytes (write_snapshot.c:188)");

namespace {

struct File {
	FILE *fp = nullptr;
	~File()
	{
		if (fp)
			fclose(fp);
	}
};

bool put_record(FILE *fp, const void *data, uint32_t size)
{
	return fwrite(&size, 4, 1, fp) == 1
			&& (size == 0 || fwrite(data, size, 1, fp) == 1)
			&& fwrite(&size, 4, 1, fp) == 1;
}

/* the label record, then the data record */
bool put_block(FILE *fp, const char label[4], const void *data, size_t size)
{
	if (size > 0xFFFFFFFFull - 8)
		return false;
	unsigned char head[8];
	const uint32_t next = (uint32_t)size + 8;
	memcpy(head, label, 4);
	memcpy(head + 4, &next, 4);
	return put_record(fp, head, 8) && put_record(fp, data, (uint32_t)size);
}

/* per-particle interleave of ncomp columns */
std::vector<double> interleave(const std::vector<std::vector<double>> &c,
		size_t n)
{
	std::vector<double> out(n * c.size());
	for (size_t i = 0; i < n; i++)
		for (size_t j = 0; j < c.size(); j++)
			out[i * c.size() + j] = c[j][i];
	return out;
}

/* find_block (read_snapshot.c:480-513): from the start of the file, the
 * data size of the block with this label, the file positioned at its data
 * record; -1 when the file has no such block */
long long find_block(FILE *fp, const char label[4])
{
	rewind(fp);
	for (;;) {
		uint32_t r0, r1, next;
		char lab[4];
		if (fread(&r0, 4, 1, fp) != 1)
			return -1;
		if (r0 != 8 || fread(lab, 4, 1, fp) != 1
				|| fread(&next, 4, 1, fp) != 1 || fread(&r1, 4, 1, fp) != 1
				|| r1 != 8 || next < 8)
			return -2;
		if (memcmp(lab, label, 4) == 0)
			return (long long)next - 8;
		if (fseek(fp, (long)next, SEEK_CUR) != 0)
			return -1;
	}
}

/* the data record of the block found by find_block, `size` bytes */
bool get_data(FILE *fp, void *out, size_t size)
{
	uint32_t r0, r1;
	return fread(&r0, 4, 1, fp) == 1 && r0 == size
			&& (size == 0 || fread(out, size, 1, fp) == 1)
			&& fread(&r1, 4, 1, fp) == 1 && r1 == size;
}

} // namespace

extern "C" int tdv_write_snapshot(tdv_context *c, const char *path,
		const tdv_snapshot_header *head, unsigned blocks)
{
	if (!c || !path || !head || (blocks & ~(TDV_BLOCK_ACC | TDV_BLOCK_GACC)))
		return tdv_fail(c, TDV_ERR_ARG, "tdv_write_snapshot: bad arguments");
	const size_t n = c->n;
	uint64_t total = 0;
	for (int t = 0; t < 6; t++)
		total += head->Npart[t];
	if (total != n || n == 0)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_write_snapshot: header Npart "
				"must add up to the particle count");
	std::vector<std::vector<double>> pos(3, std::vector<double>(n)),
			vel(3, std::vector<double>(n)), acc(3, std::vector<double>(n)),
			gacc(3, std::vector<double>(n));
	std::vector<double> mass(n);
	std::vector<uint32_t> id(n);
	double *pp[3] = { pos[0].data(), pos[1].data(), pos[2].data() };
	double *vp[3] = { vel[0].data(), vel[1].data(), vel[2].data() };
	double *ap[3] = { acc[0].data(), acc[1].data(), acc[2].data() };
	double *gp[3] = { gacc[0].data(), gacc[1].data(), gacc[2].data() };
	TDV_TRY(tdv_get_particles(c, pp, vp, mass.data(), ap, gp, nullptr,
			nullptr, nullptr, nullptr, id.data(), nullptr));

	/* write_gadget_header (write_snapshot.c:182-219), one file: Nall from
	 * the file's own counts; bytes the reference leaves unset are zero */
	tdv_snapshot_header h = *head;
	for (int t = 0; t < 6; t++) {
		h.Nall[t] = h.Npart[t];
		h.Nall_High_Word[t] = 0;
	}
	h.Num_Files = 1;
	memset(h.fill_bytes, 0, sizeof(h.fill_bytes));

	File f;
	f.fp = fopen(path, "wb");
	if (!f.fp)
		return tdv_fail(c, TDV_ERR_ARG, std::string("tdv_write_snapshot: "
				"cannot open ") + path);
	bool ok = put_block(f.fp, "HEAD", &h, sizeof(h));
	const std::vector<double> p = interleave(pos, n), v = interleave(vel, n);
	ok = ok && put_block(f.fp, "POS ", p.data(), p.size() * 8);
	ok = ok && put_block(f.fp, "VEL ", v.data(), v.size() * 8);
	ok = ok && put_block(f.fp, "ID  ", id.data(), n * 4);
	ok = ok && put_block(f.fp, "MASS", mass.data(), n * 8);
	if (blocks & TDV_BLOCK_ACC)
		ok = ok && put_block(f.fp, "ACC ", acc[0].data(), n * 8);
	if (blocks & TDV_BLOCK_GACC) {
		const std::vector<double> g = interleave(gacc, n);
		ok = ok && put_block(f.fp, "GACC", g.data(), g.size() * 8);
	}
	if (!ok || fflush(f.fp) != 0)
		return tdv_fail(c, TDV_ERR_ARG, std::string("tdv_write_snapshot: "
				"write failed on ") + path);
	return 0;
}

extern "C" int tdv_read_snapshot(tdv_context *c, const char *path,
		tdv_snapshot_header *head_out)
{
	if (!c || !path)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: bad arguments");
	/* find_files (read_snapshot.c:545-574): the name itself, or name.0,
	 * name.1, ... */
	std::vector<std::string> files;
	if (FILE *fp = fopen(path, "rb")) {
		fclose(fp);
		files.push_back(path);
	} else {
		for (int k = 0; k < 10000; k++) {
			const std::string name = std::string(path) + "." + std::to_string(k);
			FILE *g = fopen(name.c_str(), "rb");
			if (!g)
				break;
			fclose(g);
			files.push_back(name);
		}
	}
	if (files.empty())
		return tdv_fail(c, TDV_ERR_ARG, std::string("tdv_read_snapshot: "
				"cannot open ") + path);
	tdv_snapshot_header first{};
	std::vector<double> pos[3], vel[3], mass;
	std::vector<uint32_t> id;
	std::vector<int> ptype;
	bool have_mass = true;
	for (size_t k = 0; k < files.size(); k++) {
		File f;
		f.fp = fopen(files[k].c_str(), "rb");
		uint32_t rec = 0;
		/* find_endianess (read_snapshot.c:516-543): the first record is
		 * the 8-byte label record */
		if (!f.fp || fread(&rec, 4, 1, f.fp) != 1)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: unreadable "
					+ files[k]);
		if (rec != 8)
			return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: not a "
					"format-2 file in this byte order: " + files[k]);
		tdv_snapshot_header h;
		if (find_block(f.fp, "HEAD") != 256 || !get_data(f.fp, &h, 256))
			return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: header "
					"corrupted in " + files[k]);
		if (k == 0)
			first = h;
		size_t nf = 0;
		for (int t = 0; t < 6; t++)
			nf += h.Npart[t];
		std::vector<double> buf(3 * nf);
		const char *vlab[2] = { "POS ", "VEL " };
		std::vector<double> *dst[2] = { pos, vel };
		for (int b = 0; b < 2; b++) {
			if (find_block(f.fp, vlab[b]) != (long long)(24 * nf)
					|| !get_data(f.fp, buf.data(), 24 * nf))
				return tdv_fail(c, TDV_ERR_ARG, std::string("tdv_read_snapshot:"
						" block ") + vlab[b] + " missing or inconsistent in "
						+ files[k]);
			for (size_t i = 0; i < nf; i++)
				for (int d = 0; d < 3; d++)
					dst[b][d].push_back(buf[3 * i + d]);
		}
		std::vector<uint32_t> ids(nf);
		if (find_block(f.fp, "ID  ") != (long long)(4 * nf)
				|| !get_data(f.fp, ids.data(), 4 * nf))
			return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: block ID "
					"missing or inconsistent in " + files[k]);
		id.insert(id.end(), ids.begin(), ids.end());
		const long long msz = find_block(f.fp, "MASS");
		if (msz == (long long)(8 * nf)) {
			if (!get_data(f.fp, buf.data(), 8 * nf))
				return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: block "
						"MASS corrupted in " + files[k]);
			mass.insert(mass.end(), buf.begin(), buf.begin() + nf);
		} else if (msz == -1) {
			have_mass = false;             /* optional (io.h:74) */
			mass.insert(mass.end(), nf, 0.0);
		} else {
			return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: block MASS "
					"inconsistent in " + files[k]);
		}
		for (int t = 0; t < 6; t++)
			ptype.insert(ptype.end(), h.Npart[t], t);
	}
	const size_t n = id.size();
	if (n == 0)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: no particles");
	/* generate_masses_from_header (read_snapshot.c:421-446): any nonzero
	 * Massarr sets every particle's mass from its type's entry */
	bool header_mass = false;
	for (int t = 0; t < 6; t++)
		header_mass = header_mass || first.Massarr[t] != 0;
	if (header_mass)
		for (size_t i = 0; i < n; i++)
			mass[i] = first.Massarr[ptype[i]];
	else if (!have_mass)
		return tdv_fail(c, TDV_ERR_ARG, "tdv_read_snapshot: no MASS block "
				"and no header masses");
	const double *pp[3] = { pos[0].data(), pos[1].data(), pos[2].data() };
	const double *vp[3] = { vel[0].data(), vel[1].data(), vel[2].data() };
	TDV_TRY(tdv_set_particles(c, n, pp, vp, mass.data(), nullptr, nullptr,
			nullptr, id.data()));
	if (head_out)
		*head_out = first;
	return 0;
}
